"""Benchmark of the FFV-HC hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): standalone pressure Poisson on a
2-level block octree -- roots 8x8x8 of 16^3-cell blocks, level 1 inside a ball of radius L/4 at the
centre (1-cell margin: 1,128 blocks, 4.62 M cells); 7-band matrix of bcut_calc_ab_p without cut
cells (bcut.f90:608-832), outer faces Neumann except X+ Dirichlet p = 0 (the
all-Neumann 2-level operator is singular and not symmetric, so no solver reaches 1e-10 on it);
seeded uniform[-1,1] RHS (PCG64 seed 0) with the mean removed.

A step is one Poisson solve from x = 0 to relative residual ||b-Ax||/||b|| <= 1e-10 (FFVILS.h:24-47).
value = cells * steps / seconds of the timed steps (Mcell-steps/s), inputs resident in HBM.

Prints one JSON line.  --gpus N > 1 exits: every box here has one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ffv-hc_amd"))
sys.path.insert(0, ROOT)

N, G = 16, 2
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"rbgs64": 92.0, "rbgs32": 46.0}  # algorithmic bytes / cell / iteration


def build_c2(margin=1.0):
    import ffvhc
    tree = ffvhc.Tree(root=(8, 8, 8), root_length=1.0, margin=margin,
                      regions=[dict(kind="ball", a=(4.0, 4.0, 4.0), r=2.0, level=1)])
    return tree


def poisson_system(tree, rho=1.0, seed=0):
    """Padded A (7 bands) and b for the cut-free Poisson matrix with the outer BC folded in."""
    tb = tree.table()
    nb = tree.nblk
    shp = (nb, N + 2 * G, N + 2 * G, N + 2 * G)
    l = (1.0 / rho) / (tb["dx"] * tb["dx"])  # bcut.f90:773-778
    A = [np.zeros(shp) for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = l[:, None, None, None]
    A[0][:] = (-(l + l) - (l + l) - (l + l))[:, None, None, None]
    rng = np.random.default_rng(seed)
    inner = (slice(None),) + (slice(G, G + N),) * 3
    b = np.zeros(shp)
    vals = rng.uniform(-1.0, 1.0, (nb, N, N, N))
    b[inner] = vals - vals.mean()
    # matrix BC on outer faces (bca.f90): X+ Dirichlet 0, others Neumann 0
    bct = [1, 0, 1, 1, 1, 1]
    for blk in range(nb):
        for f in range(6):
            if tb["ftype"][blk, f] != 0:
                continue
            ax = f // 2
            idx = [slice(G, G + N)] * 3
            idx[2 - ax] = (G + N - 1) if f & 1 else G
            idx = (blk,) + tuple(idx)
            if bct[f] == 0:
                A[0][idx] = A[0][idx] - A[1 + f][idx]
            else:
                A[0][idx] = A[0][idx] + A[1 + f][idx]
            A[1 + f][idx] = 0.0
    return A, b, bct


def setup(tree, A, b, bct):
    import ffvhc
    ctx = ffvhc.Context(max_blocks=tree.nblk)
    params = ffvhc.case_params(bc={ffvhc.BC_P: [(t, 0.0) for t in bct]})
    cid = ctx.add_case(tree, params)
    f64 = {"A": [ctx.field(8) for _ in range(7)], "b": ctx.field(8), "x": ctx.field(8)}
    for q in range(7):
        ctx.upload(f64["A"][q], cid, A[q])
    ctx.upload(f64["b"], cid, b)
    f32 = {"A": [ctx.field(4) for _ in range(7)], "b": ctx.field(4), "x": ctx.field(4)}
    for q in range(7):
        ctx.upload(f32["A"][q], cid, A[q].astype(np.float32))
    ctx.upload(f32["b"], cid, b.astype(np.float32))
    return ctx, cid, f64, f32


def cpu_baseline(tree, A, b, bct, iters_per_solve, budget_s=15.0):
    """Oracle (C restatement, OpenMP over blocks) RBGS sweeps on the same system, host cores."""
    import ctypes
    from oracle import oracle as O
    tb = tree.table()
    tab = dict(tb)
    tab.update(g=G, ftype=tb["ftype"].ravel(), nbr=tb["nbr"].ravel(), sub=tb["sub"].ravel(), org=tb["org"].ravel())
    grid = O.make_grid(tab)
    L = O.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    L.orc_set_threads(threads)
    x = np.zeros_like(b)
    cnt, res = ctypes.c_int(), ctypes.c_double()
    args = (ctypes.byref(grid), O.ptr(x), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6([0.0] * 6),
            ctypes.c_double(1.2))
    L.orc_rbgs(*args, 2, ctypes.c_double(-1.0), ctypes.byref(cnt), ctypes.byref(res))
    sweeps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        L.orc_rbgs(*args, 5, ctypes.c_double(-1.0), ctypes.byref(cnt), ctypes.byref(res))
        sweeps += 5
    dt = time.perf_counter() - t0
    cells = tree.nblk * N ** 3
    per_sweep = dt / sweeps
    return {"value": cells / (per_sweep * iters_per_solve) / 1e6, "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d RBGS sweeps (+residual) of the same %d-block system in %.1f s on %d threads "
                      "(%.2f Mcell-sweeps/s), scaled by the %d sweeps the FP64 solve needs"
                      % (sweeps, tree.nblk, dt, threads, cells * sweeps / dt / 1e6, iters_per_solve)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--count-max", type=int, default=400000)
    ap.add_argument("--omega", type=float, default=1.2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-sweeps", type=int, default=0,
                    help="only run this many fused FP64 and FP32 RBGS iterations (for rocprofv3) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: every box in this run has one GPU; multi-GPU runs are not part of this build "
              "(the sharding design is in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    tree = build_c2()
    cells = tree.nblk * N ** 3
    A, b, bct = poisson_system(tree)
    ctx, cid, f64, f32 = setup(tree, A, b, bct)
    import ffvhc
    if args.profile_sweeps:
        for f in (f64, f32):
            ctx.fill(f["x"], 0.0)
            ms = ctx.rbgs_fixed(f["x"], f["A"], f["b"], ffvhc.BC_P, args.omega, args.profile_sweeps)
            print("profile: %d sweeps, %.2f us/sweep" % (args.profile_sweeps, ms * 1e3 / args.profile_sweeps))
        ctx.close()
        return

    # kernel time of the fused RBGS iteration (event-timed back-to-back launches)
    ctx.fill(f64["x"], 0.0)
    ctx.rbgs_fixed(f64["x"], f64["A"], f64["b"], ffvhc.BC_P, args.omega, 20)
    ms64 = ctx.rbgs_fixed(f64["x"], f64["A"], f64["b"], ffvhc.BC_P, args.omega, 200) / 200
    ctx.fill(f32["x"], 0.0)
    ctx.rbgs_fixed(f32["x"], f32["A"], f32["b"], ffvhc.BC_P, args.omega, 20)
    ms32 = ctx.rbgs_fixed(f32["x"], f32["A"], f32["b"], ffvhc.BC_P, args.omega, 200) / 200

    def solve():
        ctx.fill(f64["x"], 0.0)
        cnt, res, rc = ctx.rbgs(f64["x"], f64["A"], f64["b"], ffvhc.BC_P, args.omega, args.count_max, args.eps,
                                ncases=1)
        return int(cnt[0]), float(res[0])

    for _ in range(args.warmup):
        solve()
    ctx.sync()
    t0 = time.perf_counter()
    stats = [solve() for _ in range(args.steps)]
    ctx.sync()
    dt = time.perf_counter() - t0
    iters = stats[-1][0]
    value = cells * args.steps / dt / 1e6
    achieved = BYTES_PER_CELL["rbgs64"] * cells / (ms64 * 1e-3) / 1e9
    out = {
        "metric": "Mcell-steps/s mixed vs FP64 at equal residual; p50/p99 served-case latency",
        "value": round(value, 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform[-1,1] RHS, mean removed)",
        "config": {"workload": "standalone Poisson, 2-level octree, roots 8^3 x 16^3 cells, ball r=L/4 level 1",
                   "blocks": tree.nblk, "cells": cells, "solver": "RBGS omega=%.2f" % args.omega,
                   "eps": args.eps, "iterations_per_step": iters, "final_residual": stats[-1][1]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "rbgs_iter_kernel<double>", "kernel_us": round(ms64 * 1e3, 2),
                     "bytes_per_cell": BYTES_PER_CELL["rbgs64"]},
        "fp32_sweep": {"kernel_us": round(ms32 * 1e3, 2),
                       "achieved_GBs": round(BYTES_PER_CELL["rbgs32"] * cells / (ms32 * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tree, A, b, bct, iters, args.cpu_seconds)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
