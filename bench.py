"""Benchmark of the FFV-HC hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): standalone pressure Poisson on a
2-level block octree -- roots 8x8x8 of 16^3-cell blocks, level 1 inside a ball of radius L/4 at the
centre (1-cell margin: 1,128 blocks, 4.62 M cells); 7-band matrix of bcut_calc_ab_p without cut
cells (bcut.f90:608-832), outer faces Neumann except X+ Dirichlet p = 0 (the all-Neumann 2-level
operator is singular and not symmetric, so no solver reaches 1e-10 on it);
seeded uniform[-1,1] RHS (PCG64 seed 0) with the mean removed.

A step is one Poisson solve from x = 0 to relative residual ||b-Ax||/||b|| <= 1e-10 (FFVILS.h:24-47),
the residual always computed in FP64 with the FP64 matrix.  The FP64 reference path is the reference's
solver, Jacobi right-preconditioned BiCGSTAB (FFVILS.h:793-974; the coarse/fine interpolation makes the
octree operator non-symmetric, where Jacobi-PCG stagnates).  value = cells * steps / seconds of the
mixed-precision schedule the autotuner picks (timed on three other seeded right-hand sides of the
same matrix) among FP32-inner defect-correction solvers and FP64-recurrence BiCGSTAB variants on an
FP32-stored matrix with block red-black GS / SOR preconditioners; the FP64 path is timed beside it.

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


def build_c2(margin=1.0):
    import ffvhc
    tree = ffvhc.Tree(root=(8, 8, 8), root_length=1.0, margin=margin,
                      regions=[dict(kind="ball", a=(4.0, 4.0, 4.0), r=2.0, level=1)])
    return tree


def seeded_rhs(tree, seed):
    """Padded right-hand side: seeded uniform[-1, 1] interior values with the mean removed."""
    nb = tree.nblk
    rng = np.random.default_rng(seed)
    b = np.zeros((nb, N + 2 * G, N + 2 * G, N + 2 * G))
    vals = rng.uniform(-1.0, 1.0, (nb, N, N, N))
    b[(slice(None),) + (slice(G, G + N),) * 3] = vals - vals.mean()
    return b


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
    b = seeded_rhs(tree, seed)
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


def oracle_grid(tree):
    from oracle import oracle as O
    tb = tree.table()
    tab = dict(tb)
    tab.update(g=G, ftype=tb["ftype"].ravel(), nbr=tb["nbr"].ravel(), sub=tb["sub"].ravel(), org=tb["org"].ravel())
    return O.make_grid(tab), tab


def cpu_baseline(tree, A, b, bct, iters_per_solve, budget_s=15.0):
    """Oracle (C restatement of FFVILS::BiCGSTAB, OpenMP over blocks) on the same system, host cores:
    a bounded number of iterations, scaled by the iterations the FP64 solve needs."""
    import ctypes
    from oracle import oracle as O
    grid, _tab = oracle_grid(tree)
    L = O.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    L.orc_set_threads(threads)
    scratch = [np.zeros_like(b) for _ in range(10)]
    cnt, res = ctypes.c_int(), ctypes.c_double()

    def run(k):
        x = scratch[9]
        x[:] = 0.0
        L.orc_bicgstab(ctypes.byref(grid), O.ptr(x), O.parr(A), O.ptr(b), O.parr(scratch[:9]),
                       O.cint6(bct), O.cdbl6([0.0] * 6), ctypes.c_double(1.0), 1, k, ctypes.c_double(-1.0),
                       ctypes.byref(cnt), ctypes.byref(res))

    run(2)
    k, t0 = 4, time.perf_counter()
    run(k)
    dt = time.perf_counter() - t0
    while dt < budget_s / 4 and k < iters_per_solve:
        k *= 2
        t0 = time.perf_counter()
        run(k)
        dt = time.perf_counter() - t0
    cells = tree.nblk * N ** 3
    per_it = dt / k
    return {"value": cells / (per_it * iters_per_solve) / 1e6, "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port", "host": host_info(),
            "sample": "%d Jacobi-BiCGSTAB iterations of the same %d-block system in %.2f s on %d threads, scaled by "
                      "the %d iterations the FP64 solve to 1e-10 needs" % (k, tree.nblk, dt, threads,
                                                                           iters_per_solve)}


def host_info():
    """Host cores as the box reports them: nproc, the CPU model and the physical core count."""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info.update(model=kv.get("Model name"), sockets=kv.get("Socket(s)"),
                    cores_per_socket=kv.get("Core(s) per socket"), threads_per_core=kv.get("Thread(s) per core"))
    except Exception as e:  # lscpu missing: report what is known
        info["lscpu"] = str(e)
    return info


def cpu_time_step_baseline(steps=4, threads=16):
    """The oracle's Solver::Update (C restatement, OpenMP over blocks as FFV_THREAD=BLOCK, -O3) on
    the single-level flow past a sphere (8x4x4 blocks of 16^3, Re 100, dt 0.01, explicit U),
    pressure by Jacobi-BiCGSTAB to 1e-8 as the GPU FP64 path (the reference's RBGS at omega 1.2
    does not reach 1e-8 on this grid within 20,000 sweeps a step).  Stage times in the layout of
    Solver::times[] (Solver.cpp:3346-3413); value = cells x steps / sum of times[2..6] (UX, UY, UZ,
    P, U); per step a line in the layout of data-ils.txt (Solver::PrintILS, Solver.cpp:2531-2598:
    step, then count and residual of UX, UY, UZ, P, T)."""
    import ctypes
    import flows
    from oracle import oracle as O
    threads = min(threads, os.cpu_count() or 1)
    st = oracle_c1(1e-8, threads)
    L, grid, F, P, nb = st["L"], st["grid"], st["F"], st["P"], st["nb"]
    times, ils = [], []
    for n in range(steps):
        info = O.StepInfo()
        L.orc_update(ctypes.byref(grid), ctypes.byref(F), ctypes.byref(P), n, ctypes.byref(info))
        times.append([float(info.times[q]) for q in range(7)])
        ils.append("%010d %04d %.16f %04d %.16f %04d %.16f %04d %.16f %04d %.16f" % (
            n, 0, 0.0, 0, 0.0, 0, 0.0, info.countP, info.residualP, 0, 0.0))
    cells = nb * N ** 3
    t26 = sum(sum(t[2:7]) for t in times)
    return {"value": cells * steps / t26 / 1e6, "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": "%d time steps of the single-level flow past a sphere (%d cells, Re 100, dt 0.01, explicit U, "
                      "pressure by Jacobi-BiCGSTAB to 1e-8) on %d threads, timed as Solver::times[2..6]"
                      % (steps, cells, threads),
            "times_per_step": [{"UX": t[2], "UY": t[3], "UZ": t[4], "P": t[5], "U": t[6]} for t in times],
            "data_ils": ils, "host": host_info()}


def oracle_c1(eps, threads):
    """The oracle's state for the single-level flow past a sphere (8 x 4 x 4 blocks, Re 100, dt 0.01,
    explicit U) from the uniform flow u = (1, 0, 0): grid, fields, parameters (pressure by
    Jacobi-BiCGSTAB to eps) and the host arrays behind them."""
    import ctypes
    import flows
    from oracle import oracle as O
    tree, geo, counts = flows.single_level_sphere()
    grid, tab = oracle_grid(tree)
    L = O.lib()
    L.orc_set_threads(threads)
    nb = tree.nblk
    shp = (nb, N + 2 * G, N + 2 * G, N + 2 * G)
    c, cid, pid = geo
    c, cid, pid = c.reshape((6,) + shp), cid.reshape((6,) + shp), pid.reshape(shp)
    z = lambda: np.zeros(shp)
    a = {"ux": np.ones(shp), "uy": z(), "uz": z(), "p": z(), "b": z(), "lapp": z(), "t": z()}
    lists = {k: [z() for _ in range(n)] for k, n in (("uc", 3), ("ucp", 3), ("ud", 3), ("udp", 3), ("f", 3),
                                                    ("v", 6), ("A", 7), ("cgs", 5), ("bis", 9))}
    params = flows.flow_params(100.0, 0.01)
    bct = lambda grp: [params.bct[grp][f] for f in range(6)]
    bcv = lambda grp: [params.bcv[grp][f] for f in range(6)]
    for d, key in enumerate(("ux", "uy", "uz")):
        L.orc_ghost_fill(ctypes.byref(grid), O.ptr(a[key]), 5, O.cint6(bct(d)), O.cdbl6(bcv(d)))
    for q in range(6):
        lists["v"][q][:] = a[("ux", "uy", "uz")[q // 2]]
    F = O.Fields()
    for k in ("ux", "uy", "uz", "p", "b", "lapp"):
        setattr(F, k, O.ptr(a[k]))
    F.t0, F.pid = O.ptr(a["t"]), O.ptr(pid)
    for key, n in (("uc", 3), ("ucp", 3), ("ud", 3), ("udp", 3), ("f", 3), ("v", 6), ("A", 7), ("cgs", 5),
                   ("bis", 9)):
        arr = getattr(F, key)
        for q in range(n):
            arr[q] = O.ptr(lists[key][q])
    for q in range(6):
        F.c[q], F.cid[q] = O.ptr(c[q]), O.ptr(cid[q])
    P = O.Params()
    P.rho, P.mu, P.dt = params.rho, params.mu, params.dt
    for d in range(3):
        for f in range(6):
            P.bct_u[d][f], P.bcv_u[d][f] = params.bct[d][f], params.bcv[d][f]
    for f in range(6):
        P.bct_p[f], P.bcv_p[f] = params.bct[3][f], params.bcv[3][f]
    P.solver, P.omega, P.countMax, P.eps, P.countPre, P.ncells = 3, 1.0, 20000, eps, 1, nb * N ** 3
    return {"L": L, "grid": grid, "F": F, "P": P, "nb": nb, "tree": tree, "geo": geo, "params": params,
            "arrays": a, "lists": lists, "keep": (c, cid, pid, tab)}


def field_error_leg(steps=3, eps=1e-10, method=26, omega=1.6, threads=16):
    """Error against the FP64 oracle, per field: `steps` time steps of the single-level flow past a
    sphere (8 x 4 x 4 blocks, from the uniform flow, explicit U, pressure to eps) on the oracle (the C
    restatement of Solver::Update, FP64, Jacobi-BiCGSTAB), on the device's FP64 path (the same
    solver) and on the device's mixed-precision path (solve_mixed `method`); for each field the
    largest |device - oracle| over the interior cells, absolute and relative to the field's largest
    |value|."""
    import ctypes
    import ffvhc
    from oracle import oracle as O
    from serve import uniform_flow_init
    threads = min(threads, os.cpu_count() or 1)
    st = oracle_c1(eps, threads)
    L, grid, F0, P0, nb = st["L"], st["grid"], st["F"], st["P"], st["nb"]
    its_o = []
    for n in range(steps):
        info = O.StepInfo()
        L.orc_update(ctypes.byref(grid), ctypes.byref(F0), ctypes.byref(P0), n, ctypes.byref(info))
        its_o.append(int(info.countP))
    a, lists = st["arrays"], st["lists"]
    ref = {"ux": a["ux"], "uy": a["uy"], "uz": a["uz"], "p": a["p"], "lapp": a["lapp"]}
    for q in range(6):
        ref["v%d" % q] = lists["v"][q]
    shp = (nb, N + 2 * G, N + 2 * G, N + 2 * G)
    inner = (slice(None),) + (slice(G, G + N),) * 3
    out = {"workload": "%d time steps of the single-level flow past a sphere (8 x 4 x 4 blocks, uniform flow start, "
                       "explicit U, pressure to %g), oracle on %d host threads" % (steps, eps, threads),
           "oracle_pressure_iterations": its_o, "fields": {}}
    tree, geo, params = st["tree"], st["geo"], st["params"]
    for tag, kw in (("fp64", dict(solver=3, omega=1.0, count_max=20000, eps=eps)),
                    ("mixed", dict(solver=3, omega=omega, count_max=20000, eps=eps, mixed=True,
                                   mixed_method=method))):
        ctx = ffvhc.Context(max_blocks=tree.nblk)
        cid = ctx.add_case(tree, params, geometry=geo)
        F = ctx.step_fields()
        ctx.select([cid])
        uniform_flow_init(0, u0=1.0, noise=0.0)(ctx, cid, F)
        its = []
        for _ in range(steps):
            info, _ms, rc = ctx.step(F, ncases=1, **kw)
            if rc:
                raise RuntimeError("time step failed with status %d" % rc)
            its.append(int(info[0].countP))
        got = {"ux": F.ux, "uy": F.uy, "uz": F.uz, "p": F.p, "lapp": F.lapp}
        for q in range(6):
            got["v%d" % q] = F.v[q]
        for k, f in got.items():
            d = ctx.download(f, cid, out=np.zeros(shp))[inner]
            r = ref[k][inner]
            err = float(np.max(np.abs(d - r)))
            scale = float(np.max(np.abs(r)))
            out["fields"].setdefault(k, {})[tag] = {"max_abs": err, "rel": err / scale if scale > 0 else err}
        out[tag + "_pressure_iterations"] = its
        ctx.close()
    return out


def serve_leg(ncases, seed, budget_s, method=None, omega=1.0, rate=2000.0):
    """Served stream of small cases without a body: flow cases of uneven block counts arrive as a
    Poisson process of `rate` per second and run a number of time steps each under continuous
    batching; p50/p99 time-to-solution (queueing included).  The default rate puts every arrival
    within the first few tens of milliseconds, far above what either solver finishes, so the GPU,
    not the arrivals, sets cases/hour (at 50 per second the tuned solver kept up with the
    arrivals and cases/hour measured the arrival process).  The pressure solve is the reference's
    FP64 Jacobi-BiCGSTAB, or ffvhc_solve_mixed `method` with `omega`."""
    import ffvhc
    from serve import Request, Server, uniform_flow_init
    rng = np.random.default_rng(seed)
    bc = {ffvhc.BC_UX: [(0, 1.0)] + [(1, 0.0)] * 5,
          ffvhc.BC_P: [(1, 0.0), (0, 0.0)] + [(1, 0.0)] * 4}
    reqs, t = [], 0.0
    for i in range(ncases):
        nx, ny = int(rng.integers(2, 7)), int(rng.integers(2, 5))
        regions = []
        if rng.random() < 0.5:  # refined ball: a 2-level case
            regions = [dict(kind="ball", a=(nx / 2, ny / 2, ny / 2), r=float(rng.uniform(0.3, 0.8)), level=1)]
        tree = ffvhc.Tree(root=(nx, ny, ny), root_length=1.0, regions=regions)
        t += float(rng.exponential(1.0 / rate))
        reqs.append(Request(name="case%d" % i, arrival=t, tree=tree,
                            params=ffvhc.case_params(mu=float(1.0 / rng.uniform(50, 300)), dt=0.002, bc=bc),
                            steps=int(rng.integers(5, 21)), init=uniform_flow_init(i)))
    cap = 4 * max(r.tree.nblk for r in reqs)
    ctx = ffvhc.Context(max_blocks=cap)
    F = ctx.step_fields()
    if method is None:
        kw = dict(solver=3, omega=1.0, count_max=2000, eps=1e-8)
    else:
        kw = dict(solver=3, omega=omega, count_max=2000, eps=1e-8, mixed=True, mixed_method=method)
    srv = Server(ctx, F, kw)
    srv.run(reqs, t_limit=budget_s)
    rep = srv.report()
    rep.update(solver=("FP64 Jacobi-BiCGSTAB" if method is None else
                       "solve_mixed method %d, omega %.1f (%s)" % (method, omega, MIXED_METHODS[method])) + " to 1e-8",
               blocks_capacity=cap, blocks_per_case=[int(r.tree.nblk) for r in reqs],
               arrival_rate_per_s=round(ncases / t, 2))
    ctx.close()
    return rep


# algorithmic bytes per interior cell of the fixed stages of a time step (fused minimal forms):
# explicit U predictor 276 (implicit U: calc_abd_u 180 per component, plus its BiCGSTAB
# iterations), add_g/add_f/remove_p 164, calc_ab_p 212, corr_u 236, halo exchanges of 12 fields 144
STEP_FIXED_BYTES = 276 + 164 + 212 + 236 + 144
STEP_FIXED_BYTES_IMPLICIT_U = 3 * 180 + 164 + 212 + 236 + 144
# FP64 residual check of a mixed solve (r = b - A x in FP64, with the FP64 bands): 9 reads + 1 write
DEFECT_BYTES = 80


def step_iteration_bytes(method, uniform_fraction):
    """Bytes per cell of one Krylov iteration of the pressure solve: the reference's FP64 Jacobi-
    BiCGSTAB (method None) or solve_mixed `method`, whose class-uniform blocks read 27 rows instead
    of per-cell bands."""
    if method is None:
        by = kernel_bytes(8)
        return sum(by[k] for k in ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr"))
    u = sum(kernel_bytes_mixed(method, True).values())
    o = sum(kernel_bytes_mixed(method, False).values())
    return uniform_fraction * u + (1.0 - uniform_fraction) * o


def serve_c5_leg(ncases, seed, budget_s, method=None, omega=1.0, hbm_gib=96.0, pool_gib=16.0, max_steps=800,
                 rate=2.0):
    """Served stream of flows past a sphere: cases arrive as a Poisson process of `rate` per second
    (well above what the GPU finishes, so the queue never empties early), block counts log-uniform
    in [100, 8000], 1-3 levels (sphere and wake refinement), Re uniform in [50, 300], dt at CFL 0.3
    on the finest cells, implicit U (alpha 1), pressure to 1e-8 per step; each case runs until its
    drag coefficient changes by less than 1e-4 over 100 steps (or max_steps).  Admission: the
    context's fields take (hbm_gib - pool_gib) GiB over all block slots; the coarse operators of
    the two-level pressure solver come from a pool of pool_gib GiB.  The pressure solve is the
    reference's FP64 Jacobi-BiCGSTAB, or solve_mixed `method` with `omega`."""
    import ffvhc
    import flows
    from serve import DragStop, Request, Server, uniform_flow_init
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for i in range(ncases):
        target = float(np.exp(rng.uniform(np.log(100.0), np.log(8000.0))))
        levels = int(rng.integers(1, 4))
        tree, center = flows.served_case(target, levels)
        re = float(rng.uniform(50.0, 300.0))
        params = flows.flow_params(re, flows.cfl_dt(tree, 0.3))
        t += float(rng.exponential(1.0 / rate))

        def geometry(tree=tree, center=center):
            c, cid, pid, _counts = tree.geometry_sphere(center=center, radius=0.5, cut_id=1)
            return c, cid, pid

        reqs.append(Request(name="case%d" % i, arrival=t, tree=tree, params=params, steps=max_steps,
                            init=uniform_flow_init(i, noise=0.0), geometry=geometry, drag_stop=DragStop()))
    # device bytes per block slot: fields of the step and the solver's scratch, measured on a
    # small context after one step of a small case
    probe = ffvhc.Context(max_blocks=64)
    pf = probe.step_fields()
    ptree, pcenter = flows.served_case(32, 1)
    pc = probe.add_case(ptree, flows.flow_params(100.0, 0.01),
                        geometry=ptree.geometry_sphere(center=pcenter, radius=0.5, cut_id=1)[:3])
    probe.select([pc])
    uniform_flow_init(0, noise=0.0)(probe, pc, pf)
    ukw = dict(implicit_u=True, alpha_u=1.0, omega_u=1.0, count_max_u=1000, eps_u=1e-8)
    kw = dict(solver=3, omega=1.0, count_max=5000, eps=1e-8, **ukw)
    if method is not None:
        kw.update(omega=omega, mixed=True, mixed_method=method)
    probe.step(pf, ncases=1, **kw)
    slot_bytes = probe.device_bytes() / 64.0
    probe.close()
    cap = int((hbm_gib - pool_gib) * 2 ** 30 / slot_bytes)
    ctx = ffvhc.Context(max_blocks=cap)
    F = ctx.step_fields()
    two_level = method is not None and method >= 17
    srv = Server(ctx, F, kw, pool_bytes=int(pool_gib * 2 ** 30),
                 case_bytes=(None if two_level else (lambda r: 0)), log=log)
    srv.run(reqs, t_limit=budget_s)
    rep = srv.report()
    fin = srv.finished
    rep.update(solver=("FP64 Jacobi-BiCGSTAB" if method is None else
                       "solve_mixed method %d, omega %.1f" % (method, omega)) + " to 1e-8, implicit U",
               workload="flows past a sphere: blocks log-uniform in [100, 8000], 1-3 levels, Re U[50, 300], "
                        "CFL 0.3; a case stops when its drag coefficient changes by < 1e-4 over 100 steps "
                        "(at most %d steps)" % max_steps,
               hbm_gib=hbm_gib, pool_gib=pool_gib, slot_bytes=int(slot_bytes), block_slots=cap,
               arrival_rate_per_s=rate,
               cases_detail=[{"blocks": int(r.tree.nblk), "levels": int(r.tree.max_level + 1),
                              "re": round(1.0 / r.params.mu, 1), "steps": r.done_steps, "stop": r.stop_reason,
                              "cd": round(r.drag[-1], 5) if r.drag else None,
                              "tts_s": round(r.finished - r.arrival, 3)} for r in fin],
               unfinished_blocks=[int(r.tree.nblk) for r in srv.active + srv.waiting])
    ctx.close()
    return rep


def c4_leg(ncases=64, steps=3, seed=0):
    """Batched sweep of independent cases seeded by index: sphere D uniform in [0.5, 1.5], Re
    uniform in [50, 300], 1-3 levels, 200-2,000 blocks (log-uniform target) of 16^3 cells, implicit
    U, dt at CFL 0.3 on each case's finest cells, pressure to 1e-8.  Every case is resident in one
    context and each stage of the time step is one launch over all cases' blocks.  The autotuner
    times two batched steps of each candidate pressure schedule (the multiplicative two-level
    solver with 3, 4 or 5 SOR sweeps) and the batch then runs `steps` timed steps with the fastest."""
    import ffvhc
    import flows
    from serve import uniform_flow_init
    trees, params = [], []
    for i in range(ncases):
        rng = np.random.default_rng(1000 + i)
        d = float(rng.uniform(0.5, 1.5))
        re = float(rng.uniform(50.0, 300.0))
        levels = int(rng.integers(1, 4))
        target = float(np.exp(rng.uniform(np.log(200.0), np.log(2000.0))))
        tree, center = flows.served_case(target, levels, d=d)
        trees.append((tree, center, d))
        params.append(flows.flow_params(re, flows.cfl_dt(tree, 0.3), d=d))
    nblk = sum(t[0].nblk for t in trees)
    cells = nblk * N ** 3
    log("c4: %d cases, %d blocks, %d cells; building geometry" % (ncases, nblk, cells))
    ctx = ffvhc.Context(max_blocks=nblk)
    F = ctx.step_fields()
    ids = []
    t0 = time.perf_counter()
    for (tree, center, d), prm in zip(trees, params):
        geo = tree.geometry_sphere(center=center, radius=0.5 * d, cut_id=1)[:3]
        cid = ctx.add_case(tree, prm, geometry=geo)
        del geo
        ctx.select([cid])
        uniform_flow_init(cid, noise=0.0)(ctx, cid, F)
        ids.append(cid)
        if len(ids) % 8 == 0:
            log("c4: %d cases resident (%.1f s)" % (len(ids), time.perf_counter() - t0))
    setup_s = time.perf_counter() - t0
    log("c4: geometry and upload %.1f s" % setup_s)
    ctx.select(ids)
    ukw = dict(implicit_u=True, alpha_u=1.0, omega_u=1.0, count_max_u=1000, eps_u=1e-8)
    cands = [(25, 1.6), (26, 1.6), (27, 1.6)]
    tuned = []
    for m, om in cands:
        kw = dict(solver=3, omega=om, count_max=20000, eps=1e-8, mixed=True, mixed_method=m, **ukw)
        ctx.step(F, ncases=ncases, **kw)  # setup of this method (the step keeps it while A is unchanged)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(2):
            info, _ms, rc = ctx.step(F, ncases=ncases, **kw)
            if rc:
                raise RuntimeError("c4 step failed with status %d" % rc)
        ctx.sync()
        tuned.append((m, om, (time.perf_counter() - t1) / 2))
        log("c4 candidate method %d: %.3f s per batched step" % (m, tuned[-1][2]))
    m, om, _ = min(tuned, key=lambda r: r[2])
    kw = dict(solver=3, omega=om, count_max=20000, eps=1e-8, mixed=True, mixed_method=m, **ukw)
    ctx.step(F, ncases=ncases, **kw)
    ctx.sync()
    t1 = time.perf_counter()
    its, res = [], []
    for _ in range(steps):
        info, _ms, rc = ctx.step(F, ncases=ncases, **kw)
        if rc:
            raise RuntimeError("c4 step failed with status %d" % rc)
        its.append([int(q.countP) for q in info])
        res.append(max(float(q.residualP) for q in info))
    ctx.sync()
    dt = (time.perf_counter() - t1) / steps
    umax = float(np.max(ctx.stats(F.ux, ncases=ncases)[:, 3]))
    ctx.close()
    lv = np.bincount([t[0].max_level for t in trees], minlength=3)
    return {"workload": "batched sweep of %d cases seeded by index: sphere D U[0.5, 1.5], Re U[50, 300], 1-3 levels, "
                        "200-2,000 blocks of 16^3 (log-uniform), implicit U, CFL 0.3, pressure to 1e-8; one launch "
                        "per stage over all cases" % ncases,
            "cases": ncases, "blocks": nblk, "cells": cells, "cases_per_level_count": [int(v) for v in lv],
            "blocks_min_max": [min(t[0].nblk for t in trees), max(t[0].nblk for t in trees)],
            "mcell_steps_s": round(cells / dt / 1e6, 3), "ms_per_batched_step": round(dt * 1e3, 2),
            "pressure_iterations_mean": round(float(np.mean(its)), 1),
            "pressure_iterations_max": int(np.max(its)), "pressure_residual_max": max(res), "ux_abs_max": umax,
            "schedule": "solve_mixed method %d (multiplicative two-level, %d SOR sweeps), omega %.1f"
                        % (m, m - 22, om),
            "autotune": [{"method": c[0], "omega": c[1], "s_per_step": round(c[2], 4)} for c in tuned],
            "setup_s": round(setup_s, 1)}


def time_step_leg(case, steps, warmup, method, omega, fp64_steps=None):
    """Fractional time steps (explicit U) of a flow past a sphere, pressure to the case's eps each
    step, with the reference's FP64 Jacobi-BiCGSTAB and with solve_mixed `method`.  case: "c1" --
    single level, 8x4x4 blocks of 16^3 cells (128x64x64, h = D/8, domain 16D x 8D x 8D), sphere
    D = 1 at (4D, 4D, 4D), Re 100, dt 0.01, eps 1e-8; "c3" -- three levels refined toward the sphere
    and its wake (roots 16x8x8 of edge D, 4,048 blocks, 16.6 M cells, h = D/64 at the sphere),
    Re 300, dt at CFL 0.3 on the finest cells, eps 1e-10, implicit U (the reference's default
    SolvingMethod/U, alpha 1, FFVConfig.cpp:433-436: the explicit predictor is unstable at this
    resolution -- explicit diffusion through cut faces as short as the 0.01 cutoff -- on the
    oracle as on the device).
    The roofline of a step: algorithmic bytes = (fixed stages + pressure iterations x bytes per
    iteration + velocity iterations x FP64 BiCGSTAB bytes + FP64 residual checks) per cell x cells,
    plus the coarse inverse (FP32, n^2) read twice per pressure iteration, over the step time."""
    import ffvhc
    import flows
    from serve import uniform_flow_init
    if case == "c1":
        tree, geo, counts = flows.single_level_sphere()
        params = flows.flow_params(100.0, 0.01)
        eps = 1e-8
        desc = ("flow past a sphere, single level: 8x4x4 blocks of 16^3 (524,288 cells), Re 100, dt 0.01, "
                "pressure to 1e-8 per step")
    else:
        tree, geo, counts = flows.three_level_sphere()
        params = flows.flow_params(300.0, flows.cfl_dt(tree, 0.3))
        eps = 1e-10
        desc = ("flow past a sphere, three levels (sphere and wake refinement, SphereDivider3 rules): %d blocks "
                "of 16^3 (%d cells), Re 300, dt %.5f (CFL 0.3 on the finest cells), pressure to 1e-10 per step"
                % (tree.nblk, tree.nblk * N ** 3, params.dt))
    implicit = case != "c1"
    ukw = dict(implicit_u=True, alpha_u=1.0, omega_u=1.0, count_max_u=1000, eps_u=1e-8) if implicit else {}
    if implicit:
        desc += "; implicit U (alpha 1, FP64 Jacobi-BiCGSTAB to 1e-8)"
    cells = tree.nblk * N ** 3
    out = {"workload": desc, "blocks": tree.nblk, "levels": tree.max_level + 1, "fluid_cells": int(counts[0]),
           "solid_cells": int(counts[1])}
    # FP64 reference path: Jacobi-BiCGSTAB (the reference's RBGS does not reach 1e-8 on the
    # single-level grid within 20,000 sweeps per step)
    for tag, kw in (("fp64", dict(solver=3, omega=1.0, count_max=20000, eps=eps, **ukw)),
                    ("mixed", dict(solver=3, omega=omega, count_max=20000, eps=eps, mixed=True,
                                   mixed_method=method, **ukw))):
        nsteps = fp64_steps if (tag == "fp64" and fp64_steps) else steps
        ctx = ffvhc.Context(max_blocks=tree.nblk)
        cid = ctx.add_case(tree, params, geometry=geo)
        F = ctx.step_fields()
        ctx.select([cid])
        uniform_flow_init(0, u0=1.0, noise=0.0)(ctx, cid, F)
        for _ in range(warmup):
            ctx.step(F, ncases=1, **kw)
        ctx.sync()
        t0 = time.perf_counter()
        its, res, uits = [], [], []
        for _ in range(nsteps):
            info, _ms, rc = ctx.step(F, ncases=1, **kw)
            if rc:
                raise RuntimeError("time step failed with status %d" % rc)
            its.append(int(info[0].countP))
            res.append(float(info[0].residualP))
            uits.append(sum(int(v) for v in info[0].countU) if implicit else 0)
        ctx.sync()
        dt = time.perf_counter() - t0
        st = ctx.stats(F.ux, ncases=1)[0]
        nu, no = ctx.stencil_classes(F.A)
        m = None if tag == "fp64" else method
        per_it = step_iteration_bytes(m, nu / max(nu + no, 1))
        mean_it = float(np.mean(its))
        mean_uit = float(np.mean(uits))
        coarse = 0.0 if m is None or m < 17 else 2 * 4.0 * tree.nblk * ((tree.nblk + 3) // 4 * 4) * mean_it
        fixed = STEP_FIXED_BYTES_IMPLICIT_U if implicit else STEP_FIXED_BYTES
        step_bytes = (fixed + mean_it * per_it + mean_uit * step_iteration_bytes(None, 0.0) +
                      (0 if m is None else DEFECT_BYTES)) * cells + coarse
        gbs = step_bytes * nsteps / dt / 1e9
        out[tag] = {"steps": nsteps, "ms_per_step": round(dt / nsteps * 1e3, 3),
                    "mcell_steps_s": round(cells * nsteps / dt / 1e6, 3),
                    "poisson_iterations_mean": round(mean_it, 1), "poisson_residual_max": max(res),
                    "velocity_iterations_mean": round(mean_uit, 1), "ux_max": float(st[1]),
                    "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_step": int(step_bytes),
                                 "bytes_per_cell_fixed": fixed,
                                 "bytes_per_cell_iteration": round(per_it, 2),
                                 "class_uniform_blocks": nu, "other_blocks": no}}
        ctx.close()
    out["speedup"] = round(out["mixed"]["mcell_steps_s"] / out["fp64"]["mcell_steps_s"], 3)
    out["solver_mixed"] = "solve_mixed method %d, omega %.1f" % (method, omega)
    return out


def flow_run_leg(case, steps, method=26, omega=1.6, every=25, fp64=False):
    """The single-level ("c1") or three-level ("c3") flow past a sphere, run for `steps` steps with
    solve_mixed `method`.  Single level: 128 blocks, 524,288 cells, Re 100, dt 0.01, explicit U,
    pressure to 1e-8, u = (1, 0, 0) plus seeded N(0, 0.01) noise, seed 1.  Three levels: 4,048
    blocks, 16.6 M cells, Re 300, dt at CFL 0.3 on the finest cells, implicit U, pressure to 1e-10,
    from the uniform flow.  Every `every` steps the drag coefficient (pressure and viscous forces on the
    sphere's cut faces, PrintForceCID's sum, over 0.5 rho U^2 pi D^2 / 4), max u_x and the step's
    pressure iterations are recorded, and a progress line is logged.  fp64: the reference's FP64
    Jacobi-BiCGSTAB pressure solve instead."""
    import ffvhc
    import flows
    from serve import uniform_flow_init
    if case == "c1":
        tree, geo, counts = flows.single_level_sphere()
        params = flows.flow_params(100.0, 0.01)
        kw = dict(solver=3, omega=omega, count_max=20000, eps=1e-8, mixed=True, mixed_method=method)
        init = uniform_flow_init(1, u0=1.0, noise=0.01)
        desc = ("flow past a sphere, single level, %d blocks of 16^3 (%d cells), Re 100, dt 0.01, explicit U, "
                "pressure to 1e-8" % (tree.nblk, tree.nblk * N ** 3))
    else:
        tree, geo, counts = flows.three_level_sphere()
        params = flows.flow_params(300.0, flows.cfl_dt(tree, 0.3))
        kw = dict(solver=3, omega=omega, count_max=20000, eps=1e-10, mixed=True, mixed_method=method,
                  implicit_u=True, alpha_u=1.0, omega_u=1.0, count_max_u=1000, eps_u=1e-8)
        init = uniform_flow_init(0, u0=1.0, noise=0.0)
        desc = ("flow past a sphere, three levels (sphere and wake refinement), %d blocks of 16^3 (%d cells), "
                "Re 300, dt %.5f (CFL 0.3 on the finest cells), implicit U, pressure to 1e-10"
                % (tree.nblk, tree.nblk * N ** 3, params.dt))
    if fp64:
        kw.update(omega=1.0, mixed=False)
        kw.pop("mixed_method")
    ctx = ffvhc.Context(max_blocks=tree.nblk)
    cid = ctx.add_case(tree, params, geometry=geo)
    F = ctx.step_fields()
    ctx.select([cid])
    init(ctx, cid, F)
    area = np.pi / 4.0
    hist, its, uits = [], [], []
    ctx.sync()
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        info, _ms, rc = ctx.step(F, ncases=1, **kw)
        if rc:
            raise RuntimeError("%s step %d failed with status %d" % (case.upper(), n, rc))
        its.append(int(info[0].countP))
        uits.append(sum(int(v) for v in info[0].countU) if case != "c1" else 0)
        if n % every == 0 or n == steps:
            fp = ctx.force_p(F.p, 1, ncases=1)[0]
            fv = ctx.force_v(F.ux, F.uy, F.uz, 1, ncases=1)[0]
            cd = float(fp[0] + fv[0]) / (0.5 * params.rho * area)
            st = ctx.stats(F.ux, ncases=1)[0]
            hist.append({"step": n, "t": round(n * params.dt, 4), "cd": round(cd, 5), "ux_max": round(float(st[1]), 5),
                         "poisson_iterations": its[-1], "s": round(time.perf_counter() - t0, 2)})
            log("%s step %d: Cd %.4f, max u_x %.4f, %d pressure iterations, %.1f s" %
                (case.upper(), n, cd, st[1], its[-1], time.perf_counter() - t0))
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.close()
    return {"workload": "%s, %d steps" % (desc, steps),
            "solver": ("FP64 Jacobi-BiCGSTAB" if fp64 else "solve_mixed method %d, omega %.1f" % (method, omega)),
            "steps": steps,
            "seconds": round(dt, 2), "ms_per_step": round(dt / steps * 1e3, 3),
            "mcell_steps_s": round(tree.nblk * N ** 3 * steps / dt / 1e6, 3),
            "pressure_iterations_mean": round(float(np.mean(its)), 2), "pressure_iterations_max": int(max(its)),
            "velocity_iterations_mean": round(float(np.mean(uits)), 2), "history": hist}


MIXED_METHODS = ["FP32 RBGS in FP64 defect correction", "FP32 Jacobi-PCG in FP64 defect correction",
                 "FP32 Jacobi-BiCGSTAB in FP64 defect correction",
                 "FP32 Jacobi-BiCGSTAB with FP64 reliable updates", "the same with a seeded shadow residual",
                 "FP64 BiCGSTAB with an FP32 block red-black GS preconditioner",
                 "FP64 BiCGSTAB with an FP64 block red-black GS preconditioner",
                 "FP64 Jacobi-BiCGSTAB on an FP32-stored matrix",
                 "FP64 BiCGSTAB on an FP32-stored matrix, block red-black GS preconditioner",
                 "FP64 BiCGSTAB on an FP32-stored matrix, symmetric block red-black GS preconditioner",
                 "FP64 BiCGSTAB on an FP32-stored matrix, block red-black GS preconditioner, hashed shadow "
                 "residual"] + ["FP64 BiCGSTAB on an FP32-stored matrix, %d block red-black SOR sweep%s on a compact "
                                "FP32/FP16 preconditioner matrix" % (n, "" if n == 1 else "s") for n in range(1, 7)] + [
                     "FP64 BiCGSTAB on an FP32-stored matrix, %d block red-black SOR sweep%s on a compact "
                     "FP32/FP16 preconditioner matrix + coarse (one value per block) correction"
                     % (n, "" if n == 1 else "s") for n in range(1, 7)] + [
                     "FP64 BiCGSTAB on an FP32-stored matrix, coarse (one value per block) correction, then "
                     "%d block red-black SOR sweep%s on the remaining residual (multiplicative two-level)"
                     % (n, "" if n == 1 else "s") for n in range(1, 7)] + [
                     "FP32 BiCGSTAB on an FP32-stored matrix, coarse (one value per block) correction, then 4 block "
                     "red-black SOR sweeps on the remaining residual (method 26 on FP32 vectors), in FP64 defect "
                     "correction"]

# schedule of `bench.py --profile` (the rocprofv3 runs): FP32-stored matrix, coarse correction then
# 4 SOR sweeps, omega 1.6
PROFILE_SCHEDULE = (26, 0.5, 20000, 1.6)

# kernel names of the timed stages as rocprofv3 reports them
PROF_NAMES = {"bi_p": "bi_precond_kernel<%s, 0, %s>", "bi_q": "bi_matvec_kernel<%s, 0, %s, false, false, 0>",
              "bi_s": "bi_precond_kernel<%s, 1, %s>", "bi_t": "bi_matvec_kernel<%s, 1, %s, false, false, 0>",
              "bi_xr": "bi_xr_kernel<%s, false, false>", "rbgs": "rbgs_iter_kernel<%s, 0>",
              "cg_p": "cg_p_kernel<%s>", "cg_q": "cg_q_kernel<%s>", "cg_xr": "cg_xr_kernel<%s>"}


def prof_name(stage, w):
    """rocprofv3 name of a stage's kernel in the plain FP64 (w = 8) / FP32 (w = 4) solvers."""
    t = "float" if w == 4 else "double"
    f = PROF_NAMES[stage]
    return f % ((t,) * f.count("%s"))


def mixed_prof_name(method, stage, uniform=False):
    """rocprofv3 name of a stage's kernel in the FP64-recurrence methods 5-28 and in method 29's
    FP32 inner solve; uniform: the variant of the class-uniform blocks (methods 11-29 read their
    bands from the class rows)."""
    mode = {"bi_p": 0, "bi_q": 0, "bi_s": 1, "bi_t": 1}.get(stage)
    hr = "true" if method == 10 else "false"
    tf = lambda v: "true" if v else "false"
    reg = 1 if uniform and method >= 11 else 0  # the class-only variant (all blocks uniform)
    vt = "float" if method == 29 else "double"  # vector type
    sweeps = 4 if method == 29 else (method - 11) % 6 + 1
    # additive coarse correction (17-22): the matvec and XR add e to p_ / s_; multiplicative
    # (23-29): the sweep kernel stores p_ / s_ with e already added
    cc = 17 <= method <= 22
    if stage == "bi_xr":
        return "bi_xr_kernel<%s, %s, %s>" % (vt, hr, tf(cc))
    if stage in ("bi_q", "bi_t"):
        if method >= 11:
            return "bi_matvec_kernel<%s, %d, float, %s, %s, %d>" % (vt, mode, hr, tf(cc), reg)
        return "bi_matvec_kernel<double, %d, %s, %s, false, 0>" % (mode, "float" if method >= 7 else "double", hr)
    if method == 7:
        return "bi_precond_kernel<double, %d, float>" % mode
    if method >= 11:
        return "bi_precond_gsk_kernel<%s, %d, %d, %s, %d>" % (vt, mode, sweeps, tf(method >= 23), reg)
    tm, tz, sym = {5: ("float", "float", "false"), 6: ("double", "double", "false"),
                   8: ("float", "double", "false"), 9: ("float", "double", "true"),
                   10: ("float", "double", "false")}[method]
    return "bi_precond_gs_kernel<%d, %s, %s, %s>" % (mode, tm, tz, sym)


def pmc_traffic(name):
    """HBM bytes per launch of a kernel (rocprofv3 name) from the committed PMC summary of this
    workload (FETCH_SIZE + WRITE_SIZE, calibrated with a plain copy); None if absent."""
    path = os.path.join(ROOT, "profiles", "c2_poisson_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    v = d.get("bytes_per_launch", {}).get(name)
    return None if v is None else int(round(v))


# algorithmic bytes per cell per launch of the fused kernels: interior arrays read/written once,
# plus the ghost layer (6 x 256 per 4096 cells = 0.375 words) read or pushed once
def kernel_bytes(w):
    h = 0.375
    return {"bi_p": (6 + h) * w, "bi_q": (10 + h) * w, "bi_s": (5 + h) * w, "bi_t": (10 + h) * w, "bi_xr": 8 * w,
            "rbgs": (10 + 2 * h) * w, "cg_p": (4 + h) * w, "cg_q": (9 + h) * w, "cg_xr": 7 * w}


def kernel_bytes_mixed(method, uniform=False):
    """Bytes per cell per launch of the FP64-recurrence methods: FP64 vectors, the matrix bands in
    FP32 (ma = 4 bytes) or FP64 where the method stores them so; the GS preconditioner reads all
    seven bands, the Jacobi one the diagonal only.  uniform: every block is class-uniform, so the
    compact-record methods (11-28) read no per-cell band or record (27 rows per block)."""
    w = 4 if method == 29 else 8                    # vectors: FP32 in method 29's inner solve
    h = 0.375 * w                                   # ghost layer of p_ / s_
    mv = 4 if method >= 7 else 8                    # matvec bands
    mp = {5: 4, 6: 8, 7: 4, 8: 4, 9: 4, 10: 4}.get(method, 16 / 7)  # preconditioner bands
    if uniform and method >= 11:
        mv, mp = 0, 0
    nb = 1 if method == 7 else 7                    # (11-13: one 16-byte record per cell)
    r0 = 0 if method == 10 else w                   # hashed shadow residual: not stored
    # 17-22 add 8 B per cell (the block sum needs no extra read; the corrections are per block)
    return {"bi_p": 3 * w + nb * mp + 2 * w + h,    # r, p, q in; p, p_ out; p_ ghosts
            "bi_q": w + h + 7 * mv + r0 + w,        # p_ + ghosts, bands, r0 in; q_ out
            "bi_s": 2 * w + nb * mp + 2 * w + h,    # r, q_ in; s, s_ out; s_ ghosts
            "bi_t": w + h + 7 * mv + w + w,         # s_ + ghosts, bands, s in; t_ out
            "bi_xr": 7 * w + r0}                    # x, p_, s_, t_, s, r0 in; x, r out


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--inner-eps", type=float, default=0.0,
                    help="reliable-update factor of the mixed solve (0: chosen by the autotuner)")
    ap.add_argument("--count-max", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-cases", type=int, default=64)
    ap.add_argument("--step-steps", type=int, default=10, help="timed steps of the time-step leg (0: skip)")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: timing loops and one solve per precision")
    ap.add_argument("--time-step-only", action="store_true",
                    help="only the time steps of the flow past a sphere (method 26, omega 1.6), for rocprofv3")
    ap.add_argument("--time-step-case", default="c1", choices=("c1", "c3"),
                    help="flow of --time-step-only: c1 single level, c3 three levels")
    ap.add_argument("--c3-steps", type=int, default=5, help="timed mixed steps of the three-level leg (0: skip)")
    ap.add_argument("--c5-cases", type=int, default=8,
                    help="cases of the drag-stopped served stream (block counts log-uniform in [100, 8000]; 0: skip)")
    ap.add_argument("--c5-seconds", type=float, default=100.0, help="wall budget of the served stream")
    ap.add_argument("--c5-max-steps", type=int, default=2000, help="most steps a served case runs")
    ap.add_argument("--c5-only", action="store_true", help="only the drag-stopped served stream")
    ap.add_argument("--c4-only", action="store_true", help="only the batched sweep of 64 cases")
    ap.add_argument("--c4-cases", type=int, default=64)
    ap.add_argument("--c3-run", type=int, default=0,
                    help="only this many steps of the three-level flow past a sphere (mixed solver), drag history")
    ap.add_argument("--c1-run", type=int, default=0,
                    help="only this many steps of the single-level flow past a sphere (mixed solver), drag history")
    ap.add_argument("--run-fp64", action="store_true", help="--c1-run / --c3-run with the FP64 pressure solve")
    ap.add_argument("--field-errors", type=int, default=0,
                    help="only the per-field error of the FP64 and mixed device paths against the oracle over this "
                         "many steps of the single-level flow past a sphere")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU only; multi-GPU runs are not implemented (the "
              "block octree would shard by contiguous Hilbert-ordered block ranges per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.c4_only:
        print(json.dumps(c4_leg(args.c4_cases)))
        return
    if args.c3_run > 0:
        print(json.dumps(flow_run_leg("c3", args.c3_run, fp64=args.run_fp64)))
        return
    if args.field_errors > 0:
        print(json.dumps(field_error_leg(args.field_errors)))
        return
    if args.c1_run > 0:
        print(json.dumps(flow_run_leg("c1", args.c1_run, fp64=args.run_fp64)))
        return
    if args.c5_only:
        rep = serve_c5_leg(max(args.c5_cases, 1), 11, args.c5_seconds, method=26, omega=1.6,
                           max_steps=args.c5_max_steps)
        print(json.dumps(rep))
        return
    if args.time_step_only:
        print(json.dumps(time_step_leg(args.time_step_case, max(args.step_steps, 1), 2, 26, 1.6)))
        return
    import ffvhc
    tree = build_c2()
    cells = tree.nblk * N ** 3
    A, b, bct = poisson_system(tree)
    ctx, cid, f64, f32 = setup(tree, A, b, bct)
    # blocks whose bands are uniform per stencil position class (no cut cells here: all of them)
    n_uniform, n_other = ctx.stencil_classes(f64["A"])
    uniform = n_other == 0
    log("stencil classes: %d class-uniform blocks, %d others" % (n_uniform, n_other))
    P = ffvhc.BC_P

    log("%d blocks, %d cells resident; timing kernels" % (tree.nblk, cells))
    # per-kernel device times (HIP events on the library's stream)
    nit = 20 if args.profile else 100
    kt = {}
    for tag, f, w in (("64", f64, 8), ("32", f32, 4)):
        ctx.fill(f["x"], 0.0)
        ctx.bicgstab_timing(f["x"], f["A"], f["b"], 1.0, 5)
        ms = ctx.bicgstab_timing(f["x"], f["A"], f["b"], 1.0, nit)[:5] / nit
        ctx.fill(f["x"], 0.0)
        ctx.rbgs_fixed(f["x"], f["A"], f["b"], P, 1.2, 5)
        mr = ctx.rbgs_fixed(f["x"], f["A"], f["b"], P, 1.2, nit) / nit
        ctx.fill(f["x"], 0.0)
        ctx.cg_timing(f["x"], f["A"], f["b"], 1.0, 5)
        mc = ctx.cg_timing(f["x"], f["A"], f["b"], 1.0, nit)[:3] / nit
        by = kernel_bytes(w)
        names = ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr")
        stages = list(zip(names, ms)) + [("rbgs", mr)] + list(zip(("cg_p", "cg_q", "cg_xr"), mc))
        kt[tag] = {name: {"us": round(t * 1e3, 2), "GBs": round(by[name] * cells / (t * 1e-3) / 1e9, 1)}
                   for name, t in stages}
        log("fp%s kernels: %s" % (tag, json.dumps(kt[tag])))

    # schedule of the mixed-precision solve: picked by the autotuner on this system (untimed)
    import autotune
    if args.profile:
        sched, tuned = PROFILE_SCHEDULE, []
    elif args.inner_eps > 0:
        sched = (1, args.inner_eps, 500, 1.0)
        tuned = []
    else:
        # tuned on three other seeded right-hand sides of the same matrix, not on the timed ones
        # (on one, schedules within a few percent of each other swap places from seed to seed)
        log("autotuning the mixed-precision schedule (right-hand side seeds 1-3)")
        b_tune = [ctx.field(8) for _ in range(3)]
        for k, bf in enumerate(b_tune):
            ctx.upload(bf, cid, seeded_rhs(tree, 1 + k))
        # best of two runs per candidate and right-hand side: the second sees the launch batches a
        # repeated solve uses
        sched, tuned = autotune.tune_mixed(ctx, f64["x"], f64["A"], b_tune, P, args.eps, repeats=2)
        for bf in b_tune:
            ctx.free(bf)
        log("schedule %s; candidates %s" % (sched, [(c, round(t, 4), it, o) for c, t, it, o, r in tuned]))

    if sched[0] >= 5:  # stage times of the tuned method's own kernels
        # the schedule converges in about 40 iterations, so its stages are timed over runs of 20
        # iterations from x = 0: iterations past convergence (vanishing residuals) measured up to
        # 38% faster per kernel on the same bytes, which would flatter every figure below
        nshort, nrep = 20, (1 if args.profile else 5)
        ctx.fill(f64["x"], 0.0)
        ctx.bicgstab_timing(f64["x"], f64["A"], f64["b"], sched[3], 5, method=sched[0])
        msall = np.zeros(7)
        for _ in range(nrep):
            ctx.fill(f64["x"], 0.0)
            msall += ctx.bicgstab_timing(f64["x"], f64["A"], f64["b"], sched[3], nshort, method=sched[0])
        msall /= nshort * nrep
        ms = msall[:5]
        by = kernel_bytes_mixed(sched[0], uniform)
        kt["mixed"] = {name: {"us": round(t * 1e3, 2), "GBs": round(by[name] * cells / (t * 1e-3) / 1e9, 1),
                              "frac": round(by[name] * cells / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "kernel": mixed_prof_name(sched[0], name, uniform)}
                       for name, t in zip(("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr"), ms)}
        # per iteration: the coarse applications, and what the timing run puts between the stages
        # (an XR finalisation launch, which the solve itself forms inside the next coarse
        # application, and the event intervals of the stages the fused methods leave empty)
        kt["mixed_other_us"] = {"timing_run_finalisation_and_gaps": round(msall[5] * 1e3, 2),
                                "coarse_corrections": round(msall[6] * 1e3, 2)}
        log("method %d kernels: %s" % (sched[0], json.dumps(kt["mixed"])))
    if args.profile:
        # the other sweep counts the autotuner may pick (their sweep kernels appear in the trace
        # and counter summaries too)
        for m in (23, 24, 25, 27, 28):
            ctx.fill(f64["x"], 0.0)
            ctx.bicgstab_timing(f64["x"], f64["A"], f64["b"], PROFILE_SCHEDULE[3], 5, method=m)

    def solve64(bf):
        ctx.fill(f64["x"], 0.0)
        cnt, res, rc = ctx.bicgstab(f64["x"], f64["A"], bf, P, 1.0, 1, args.count_max, args.eps, ncases=1)
        return int(cnt[0]), float(res[0]), 1

    def solve_mixed(bf):
        ctx.fill(f64["x"], 0.0)
        cnt, res, outer, rc = ctx.solve_mixed(sched[0], f64["x"], f64["A"], bf, P, sched[3], sched[2],
                                              args.eps, sched[1], ncases=1)
        return int(cnt[0]), float(res[0]), int(outer[0])

    timings = {}
    steps = 1 if args.profile else args.steps
    warm = 0 if args.profile else args.warmup
    # every solve (warm-up and timed) gets its own seeded right-hand side (seeds 100, 101, ...),
    # uploaded before the timed region; both paths solve the same sequence
    if args.profile:
        rhs = [f64["b"]]
    else:
        rhs = []
        for i in range(warm + steps):
            bf = ctx.field(8)
            ctx.upload(bf, cid, seeded_rhs(tree, 100 + i))
            rhs.append(bf)
    for tag, fn in (("fp64", solve64), ("mixed", solve_mixed)):
        for i in range(warm):
            fn(rhs[i])
        ctx.sync()
        t0 = time.perf_counter()
        st = []
        for i in range(steps):
            st.append(fn(rhs[warm + i]))
            log("%s solve: %d iterations, residual %.3e, %d corrections" % (tag, st[-1][0], st[-1][1], st[-1][2]))
        ctx.sync()
        dt = time.perf_counter() - t0
        timings[tag] = {"seconds": dt, "iters": round(float(np.mean([q[0] for q in st])), 2),
                        "residual": max(q[1] for q in st), "outer": st[-1][2],
                        "mcell_steps_s": cells * steps / dt / 1e6, "ms_per_step": dt / steps * 1e3}
    precision_only = None
    if not args.profile:
        # what FP32 matrix storage alone buys: the same FP64 BiCGSTAB with one block red-black GS
        # sweep as preconditioner, the matrix in FP64 (solve_mixed method 6) and in FP32 (method 8,
        # FP64 arithmetic); this system's coefficients are FP32-exact, so both take the same
        # iterations and the time ratio is the storage's alone (per-cell bands, no class rows)
        pair = {}
        for m in (6, 8):
            def solve_m(bf, m=m):
                ctx.fill(f64["x"], 0.0)
                cnt, res, _o, _rc = ctx.solve_mixed(m, f64["x"], f64["A"], bf, P, 1.2, args.count_max, args.eps,
                                                    0.5, ncases=1)
                return int(cnt[0]), float(res[0])
            solve_m(rhs[0])
            ctx.sync()
            t0 = time.perf_counter()
            q = solve_m(rhs[warm])
            ctx.sync()
            pair[m] = ((time.perf_counter() - t0) * 1e3, q[0], q[1])
            log("method %d: %d iterations, %.2f ms" % (m, q[0], pair[m][0]))
        precision_only = {"ms_fp64_matrix": round(pair[6][0], 2), "ms_fp32_matrix": round(pair[8][0], 2),
                          "iterations": [pair[6][1], pair[8][1]],
                          "residuals": [pair[6][2], pair[8][2]],
                          "speedup": round(pair[6][0] / pair[8][0], 3),
                          "solver": "FP64 BiCGSTAB, one block red-black GS sweep (omega 1.2) as preconditioner, "
                                    "matrix stored in FP64 (solve_mixed 6) vs FP32 (solve_mixed 8)"}
    if args.profile:
        # calibration for HBM counters: plain field copies of known size (interior + one ghost layer)
        for w in (8, 4):
            tmp, src = ctx.field(w), (f64 if w == 8 else f32)["b"]
            for _ in range(10):
                ctx.copy(tmp, src)
            ctx.free(tmp)
        ctx.sync()
        copy_bytes = {w: 2 * w * tree.nblk * (N ** 3 + 6 * N * N) for w in (8, 4)}
        print(json.dumps({"profile": True, "kernels": kt, "timings": timings, "copy_bytes": copy_bytes}))
        ctx.close()
        return
    mix, d64 = timings["mixed"], timings["fp64"]
    # dominant kernel of the headline (mixed) path: the slowest per-iteration kernel of the
    # schedule's inner solver
    stage_set = {0: ("32", ("rbgs",)), 1: ("32", ("cg_p", "cg_q", "cg_xr")), 2: ("32", ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr")),
                 3: ("32", ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr")), 4: ("32", ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr"))}
    if "mixed" in kt:
        dtag, dstages = "mixed", ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr")
    else:
        dtag, dstages = stage_set.get(sched[0], ("64", ("bi_p", "bi_q", "bi_s", "bi_t", "bi_xr")))
    dname = max(dstages, key=lambda k: kt[dtag][k]["us"])
    dom = kt[dtag][dname]
    if dtag == "mixed":
        dkernel, dbytes = dom["kernel"], kernel_bytes_mixed(sched[0], uniform)[dname]
    else:
        dw = 4 if dtag == "32" else 8
        dkernel, dbytes = prof_name(dname, dw), kernel_bytes(dw)[dname]
    out = {
        "metric": "Mcell-steps/s mixed vs FP64 at equal residual; p50/p99 served-case latency",
        "value": round(mix["mcell_steps_s"], 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mix["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 inner / f64 defect correction" if sched[0] <= 4 or sched[0] == 29 else
                  "f64 vectors and arithmetic / f32 matrix storage" if sched[0] >= 7 else
                  "f64 / f32 preconditioner" if sched[0] == 5 else "f64"),
        "data": "synthetic (seeded uniform[-1,1] RHS, mean removed)",
        "config": {"workload": "standalone Poisson, 2-level octree: roots 8^3 of 16^3-cell blocks, level 1 in a "
                               "ball of radius L/4 (1-cell margin); outer faces Neumann except X+ Dirichlet p = 0 "
                               "(the all-Neumann two-level operator is singular and not symmetric, so no solver "
                               "reaches 1e-10 on it); every timed solve from x = 0 with its own seeded RHS",
                   "blocks": tree.nblk, "cells": cells,
                   "stencil_classes": {"uniform_blocks": n_uniform, "other_blocks": n_other},
                   "solver": "FP64 reference path: Jacobi-BiCGSTAB (PreCount 1, omega 1), the reference's "
                             "pressure solver for the octree (Jacobi-CG stagnates on the non-symmetric coarse/fine "
                             "operator); mixed: the autotuned schedule below, accepted on the FP64 residual",
                   "eps": args.eps},
        "fp64": {"value": round(d64["mcell_steps_s"], 3), "ms_per_step": round(d64["ms_per_step"], 3),
                 "iterations": d64["iters"], "residual": d64["residual"]},
        "mixed": {"iterations_inner": mix["iters"], "corrections": mix["outer"], "residual": mix["residual"],
                  "ms_per_step": round(mix["ms_per_step"], 3),
                  "schedule": {"method": MIXED_METHODS[sched[0]], "inner_eps_or_update_factor": sched[1],
                               "inner_cap": sched[2], "omega": sched[3]},
                  "autotune": [{"method": MIXED_METHODS[c[0]], "factor": c[1], "cap": c[2], "omega": c[3],
                                "seconds_3_rhs": (None if t == float("inf") else round(t, 4)), "inner_max": it,
                                "corrections": o} for c, t, it, o, r in tuned]},
        "speedup_mixed_vs_fp64": round(mix["mcell_steps_s"] / d64["mcell_steps_s"], 3),
        # where the speedup comes from: fewer iterations (the preconditioner: algorithm) and a cheaper
        # iteration (fused kernels, FP32 matrix storage, setup included in the mixed time)
        "attribution": {"iterations_fp64": d64["iters"], "iterations_mixed": mix["iters"],
                        "iteration_ratio": round(d64["iters"] / max(mix["iters"], 1e-9), 2),
                        "ms_per_iteration_fp64": round(d64["ms_per_step"] / max(d64["iters"], 1e-9), 4),
                        "ms_per_iteration_mixed_incl_setup": round(mix["ms_per_step"] / max(mix["iters"], 1e-9), 4),
                        "note": "every block of this system is class-uniform, so the mixed path reads its "
                                "matrix from 27 rows per block: FP32 storage saves almost no bytes here; the "
                                "speedup is the two-level preconditioner's (iterations) and the fused kernels'; "
                                "precision_only times FP32 matrix storage alone on per-cell bands",
                        "precision_only": precision_only},
        "roofline": {"bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(dkernel),
                     "algorithmic_bytes": int(dbytes * cells), "kernel": dkernel, "kernel_us": dom["us"],
                     "bytes_per_cell": dbytes},
        "kernels": kt,
    }
    if args.c5_cases > 0:
        log("served stream of %d flows past a sphere, drag-stopped" % args.c5_cases)
        out["serving"] = serve_c5_leg(args.c5_cases, 11, args.c5_seconds, method=sched[0], omega=sched[3],
                                      max_steps=args.c5_max_steps)
        log("serving: " + json.dumps({k: v for k, v in out["serving"].items() if k != "cases_detail"}))
    if args.serve_cases > 0:
        log("served stream of %d small cases" % args.serve_cases)
        out["serving_small"] = serve_leg(args.serve_cases, 7, 60.0, method=sched[0], omega=sched[3])
        log("serving (small cases): " +
            json.dumps({k: v for k, v in out["serving_small"].items() if k != "blocks_per_case"}))
        out["serving_small_fp64"] = serve_leg(args.serve_cases, 7, 60.0)
        log("serving (small cases, FP64 reference solver): " +
            json.dumps({k: v for k, v in out["serving_small_fp64"].items() if k != "blocks_per_case"}))
    if args.step_steps > 0:
        log("time steps of the flow past a sphere")
        out["time_step"] = time_step_leg("c1", args.step_steps, 2, sched[0], sched[3])
        log("time step: " + json.dumps(out["time_step"]))
    if args.c3_steps > 0:
        log("time steps of the three-level flow past a sphere")
        out["time_step_c3"] = time_step_leg("c3", args.c3_steps, 1, sched[0], sched[3], fp64_steps=2)
        log("time step c3: " + json.dumps(out["time_step_c3"]))
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(tree, A, b, bct, d64["iters"], args.cpu_seconds)
        log("cpu baseline of the time step")
        out["cpu_baseline_time_step"] = cpu_time_step_baseline()
        log("cpu time step: %.3f Mcell-steps/s" % out["cpu_baseline_time_step"]["value"])
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
