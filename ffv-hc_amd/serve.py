"""Continuous-batching request loop for flow cases on one GPU context.

Cases arrive over time with uneven block counts.  HBM is budgeted in two parts: the context's
fields are allocated once over all its block slots (a slot costs the same bytes whatever case
holds it), and what is left of the budget is a pool for what a case allocates beside its slots --
the dense coarse operator of the two-level pressure solvers (FP64 and FP32 copies, 12 n^2 bytes for
n blocks).  A case is admitted when the context has a contiguous run of free slots for it and its
pool bytes fit in the pool beside those of the cases already admitted.  All admitted cases advance together: one ffvhc_step call per loop iteration runs every
stage of every active case in one launch per stage, each case with its own Poisson convergence and
step counter.  A case leaves when it has run its steps or, with a drag stopping rule, when its
drag coefficient changed by less than the tolerance over the last `window` steps (the drag is the
x force on the sphere's cut faces, pressure plus viscous, as Solver::PrintForceCID sums it,
Solver.cpp:2994-3160; read every step through the allocation-free force entry point); its slots
are reused by the next admission.  Waiting cases are tried in arrival order at every step; a
later, smaller case may start while an earlier one still waits (backfill).

Time-to-solution of a case = completion time - arrival time (wall clock).  The per-case results
do not depend on the mix: every kernel works per block and every reduction per case in a fixed
order, so a case's fields (and so its drag history and its stopping step) are bit-identical alone
and in any mix.
"""
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np


@dataclass
class Request:
    name: str
    arrival: float            # seconds after the loop starts
    tree: object              # ffvhc.Tree
    params: object            # ffvhc.CaseParams
    steps: int                # time steps to run (with drag_stop: the most it may run)
    init: Callable            # init(ctx, case_id, fields) uploads the initial condition
    geometry: Optional[object] = None  # (c, cid, pid), or a callable that returns it at admission
    result: Optional[Callable] = None  # result(ctx, case_id, fields) before the case is removed
    # drag stopping rule: DragStop(window, tol, area) -- stop when |Cd(n) - Cd(n - window)| < tol
    drag_stop: Optional["DragStop"] = None
    # filled by the loop
    admitted: float = -1.0
    finished: float = -1.0
    case_id: int = -1
    done_steps: int = 0
    bytes: int = 0
    stop_reason: str = ""
    poisson_iters: List[int] = field(default_factory=list)
    drag: List[float] = field(default_factory=list)


@dataclass
class DragStop:
    window: int = 100          # steps over which the drag coefficient must settle
    tol: float = 1e-4          # largest change of Cd over the window
    area: float = 0.7853981633974483  # frontal area of the body (pi D^2 / 4 for D = 1)
    u0: float = 1.0            # free-stream speed
    cut_id: int = 1            # boundary id of the body's cut faces


def drag_settled(cd, window, tol):
    """True once the drag history cd has `window` steps behind its last value and changed by less
    than tol over them."""
    return len(cd) > window and abs(cd[-1] - cd[-1 - window]) < tol


def uniform_flow_init(seed, u0=1.0, noise=0.01, g=2):
    """Initial condition of a flow case: u = (u0, 0, 0) + seeded N(0, noise) on every component,
    p = 0, face velocities from u, ghost layers filled (the Solver's initial fills)."""

    def init(ctx, cid, F):
        import ffvhc
        nb = ctx.cases[cid].nblk
        shp = (nb, 16 + 2 * g, 16 + 2 * g, 16 + 2 * g)
        rng = np.random.default_rng(seed)
        u = [u0 + noise * rng.standard_normal(shp), noise * rng.standard_normal(shp),
             noise * rng.standard_normal(shp)]
        for d, f in enumerate((F.ux, F.uy, F.uz)):
            ctx.upload(f, cid, u[d], g=g)
            ctx.ghost_fill(f, ffvhc.BC_UX + d, 5)
        ctx.upload(F.p, cid, np.zeros(shp), g=g)
        ctx.ghost_fill(F.p, ffvhc.BC_P, 1)
        for q in range(6):
            ctx.upload(F.v[q], cid, u[q // 2], g=g)

    return init


def bytes_per_block(nfields64, nfields32=0):
    """Device bytes one block slot costs: interior 16^3 plus the one-layer halo per field."""
    per = 16 ** 3 + 6 * 16 * 16
    return per * (8 * nfields64 + 4 * nfields32)


def coarse_bytes(nblk):
    """Device bytes of a case's dense coarse operator (FP64 in place + FP32 copy, rows padded to 4)."""
    return 8 * nblk * nblk + 4 * nblk * ((nblk + 3) // 4 * 4)


class Server:
    """Admit, interleave and retire cases on one context (continuous batching).

    pool_bytes (optional): the device bytes cases may allocate beside their slots; case_bytes(r)
    is what case r takes from it (default: its coarse operator, coarse_bytes(blocks))."""

    def __init__(self, ctx, fields, step_kwargs, clock=time.perf_counter, pool_bytes=None, case_bytes=None,
                 log=None):
        self.log = log or (lambda msg: None)
        self.ctx = ctx
        self.F = fields
        self.step_kwargs = dict(step_kwargs)
        self.clock = clock
        self.hbm_budget = pool_bytes
        self._case_bytes = case_bytes or (lambda r: coarse_bytes(r.tree.nblk))
        self.bytes_in_use = 0
        self.peak_bytes = 0
        self.waiting: List[Request] = []
        self.active: List[Request] = []
        self.finished: List[Request] = []
        self.steps_run = 0
        self.cell_steps = 0

    def case_bytes(self, r):
        return self._case_bytes(r)

    def _admit(self, now):
        still = []
        for r in self.waiting:
            need = self.case_bytes(r)
            if self.hbm_budget is not None and self.bytes_in_use + need > self.hbm_budget:
                if need > self.hbm_budget:
                    raise RuntimeError("case %s needs %d pool bytes, more than the whole pool" % (r.name, need))
                still.append(r)
                continue
            geo = r.geometry() if callable(r.geometry) else r.geometry
            try:
                cid = self.ctx.add_case(r.tree, r.params, geometry=geo)
            except Exception as e:  # status -3: no contiguous room yet, keep waiting (FIFO order kept)
                if getattr(e, "code", None) != -3:
                    raise
                still.append(r)
                continue
            del geo
            r.case_id = cid
            r.admitted = now
            r.bytes = need
            self.bytes_in_use += need
            self.peak_bytes = max(self.peak_bytes, self.bytes_in_use)
            self.ctx.select([cid])
            r.init(self.ctx, cid, self.F)
            self.active.append(r)
            self.log("t=%.2f admitted %s (%d blocks), %d active" % (now, r.name, r.tree.nblk, len(self.active)))
        self.waiting = still

    def _drag(self, reqs):
        """Drag coefficients of the active cases with a drag rule (one force call per kind over all
        of them: the pressure part through the allocation-free entry point)."""
        out = {}
        by_tag = {}
        for r in reqs:
            by_tag.setdefault(r.drag_stop.cut_id, []).append(r)
        for cut_id, rs in by_tag.items():
            self.ctx.select([r.case_id for r in rs])
            fp = self.ctx.force_p_noalloc(self.F.p, cut_id, ncases=len(rs))
            fv = self.ctx.force_v(self.F.ux, self.F.uy, self.F.uz, cut_id, ncases=len(rs))
            for r, a, b in zip(rs, fp, fv):
                rho = r.params.rho
                ds = r.drag_stop
                out[id(r)] = float(a[0] + b[0]) / (0.5 * rho * ds.u0 * ds.u0 * ds.area)
        return out

    def run(self, requests, t_limit=None):
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = self.clock()
        while pending or self.waiting or self.active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                self.waiting.append(pending.pop(0))
            if self.waiting:
                self._admit(now)
            if not self.active:
                if pending:  # idle until the next arrival
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                    continue
                if self.waiting:
                    raise RuntimeError("a waiting case can never fit the context")
                break
            ids = [r.case_id for r in self.active]
            self.ctx.select(ids)
            info, _ms, rc = self.ctx.step(self.F, ncases=len(ids), **self.step_kwargs)
            if rc:
                raise RuntimeError("time step failed with status %d" % rc)
            self.steps_run += 1
            if self.steps_run % 100 == 0:
                self.log("t=%.2f %d loop steps, %d active, %d waiting" % (self.clock() - t0, self.steps_run,
                                                                         len(self.active), len(self.waiting)))
            dragged = [r for r in self.active if r.drag_stop is not None]
            cds = self._drag(dragged) if dragged else {}
            now = self.clock() - t0
            keep = []
            for r, inf in zip(self.active, info):
                r.done_steps += 1
                r.poisson_iters.append(int(inf.countP))
                self.cell_steps += r.tree.nblk * 16 ** 3
                stop = ""
                if r.drag_stop is not None:
                    r.drag.append(cds[id(r)])
                    if drag_settled(r.drag, r.drag_stop.window, r.drag_stop.tol):
                        stop = "drag settled"
                if not stop and r.done_steps >= r.steps:
                    stop = "step limit"
                if stop:
                    r.finished = now
                    r.stop_reason = stop
                    if r.result is not None:
                        self.ctx.select([r.case_id])
                        r.result(self.ctx, r.case_id, self.F)
                    self.ctx.remove_case(r.case_id)
                    self.bytes_in_use -= r.bytes
                    self.finished.append(r)
                    self.log("t=%.2f finished %s after %d steps (%s)" % (now, r.name, r.done_steps, stop))
                else:
                    keep.append(r)
            self.active = keep
            if t_limit is not None and now > t_limit:
                break
        self.ctx.select(None)
        self.wall = self.clock() - t0
        return self.finished

    def report(self):
        tts = np.array([r.finished - r.arrival for r in self.finished])
        out = {"cases": len(self.finished), "wall_s": self.wall, "loop_steps": self.steps_run,
               "mcell_steps_s": self.cell_steps / max(self.wall, 1e-12) / 1e6,
               "unfinished": len(self.active) + len(self.waiting),
               "stopped_by_drag": sum(1 for r in self.finished if r.stop_reason == "drag settled"),
               "stopped_by_step_limit": sum(1 for r in self.finished if r.stop_reason == "step limit")}
        if self.hbm_budget is not None:
            out.update(pool_bytes=int(self.hbm_budget), peak_pool_bytes=int(self.peak_bytes))
        if tts.size:
            out.update(p50_s=float(np.percentile(tts, 50)), p99_s=float(np.percentile(tts, 99)),
                       cases_per_hour=len(self.finished) / max(self.wall, 1e-12) * 3600.0)
        return out
