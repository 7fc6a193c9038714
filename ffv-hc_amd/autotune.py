"""Autotuner for the solver schedule of a resident case set.

The parameters that decide the throughput of a Poisson solve on the device are schedule
parameters, not numerics of the result: the mixed-precision method and its reliable-update factor
(how far the FP32 recurrence runs before the FP64 residual re-anchors it), the inner tolerance
and iteration cap of a defect-correction solve.  Every candidate solves the same system from the
same start to the same FP64 relative residual, so the candidates are interchangeable; the tuner
times each on the device and keeps the fastest that reached the tolerance.
"""
import time

import numpy as np


DEFAULT_CANDIDATES = (
    # (method, inner_eps or reliable-update factor, inner iteration cap, omega); see
    # ffvhc_solve_mixed for the methods
    (1, 1e-2, 500, 1.0),                 # FP32 Jacobi-PCG in FP64 defect correction
    (8, 0.5, 20000, 1.2),                # FP64 recurrence, FP32 matrix, one block GS sweep
    (15, 0.5, 20000, 1.6),               # 5 block SOR sweeps on the compact preconditioner matrix
    (19, 0.5, 20000, 1.6),               # 3 sweeps + additive coarse correction
    # coarse correction first, then 2-6 sweeps on the remaining residual (multiplicative)
    (24, 0.5, 20000, 1.6),
    (25, 0.5, 20000, 1.4),
    (25, 0.5, 20000, 1.6),
    (26, 0.5, 20000, 1.6),
    (27, 0.5, 20000, 1.6),
    (28, 0.5, 20000, 1.6),
    # method 26's preconditioner on FP32 vectors inside FP64 defect correction: inner tolerance,
    # inner cap per correction
    (29, 1e-4, 200, 1.6),
    (29, 1e-5, 200, 1.6),
)


def tune_mixed(ctx, x, A, b, group, eps, candidates=DEFAULT_CANDIDATES, repeats=1, ncases=1, x0=None):
    """Time every candidate mixed-precision schedule on the resident system; returns
    (best candidate, list of (candidate, seconds, inner iterations, corrections, residual)).
    b: one right-hand side field or a list of them; a candidate's time is the sum over the list
    of its best of `repeats` solves (iteration counts move by a few from one right-hand side to
    the next, enough to reorder schedules that tie on one), its iterations the largest count."""
    results = []
    rhs = list(b) if isinstance(b, (list, tuple)) else [b]
    for cand in candidates:
        method, ieps, cap, omega = cand
        total, ok, its, outs, resmax = 0.0, True, 0, 0, 0.0
        for bf in rhs:
            best_t = None
            for _ in range(repeats):
                if x0 is None:
                    ctx.fill(x, 0.0)
                else:
                    x0()
                ctx.sync()
                t0 = time.perf_counter()
                try:
                    cnt, res, outer, rc = ctx.solve_mixed(method, x, A, bf, group, omega, cap, eps, ieps,
                                                          ncases=ncases)
                except RuntimeError:  # a schedule the solver rejects is skipped
                    cnt, res, outer, rc = np.zeros(1, int), np.ones(1), np.zeros(1, int), 1
                ctx.sync()
                dt = time.perf_counter() - t0
                best_t = dt if best_t is None else min(best_t, dt)
            ok = ok and rc == 0 and bool(np.all(res <= eps))
            total += best_t
            its, outs, resmax = max(its, int(np.max(cnt))), max(outs, int(np.max(outer))), max(resmax, float(np.max(res)))
        results.append((cand, total if ok else float("inf"), its, outs, resmax))
    best = min(results, key=lambda r: r[1])
    if not np.isfinite(best[1]):
        raise RuntimeError("no mixed-precision schedule reached the tolerance")
    return best[0], results
