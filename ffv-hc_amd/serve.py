"""Continuous-batching request loop for flow cases on one GPU context.

Cases arrive over time with uneven block counts.  A case is admitted when the context has a
contiguous run of free block slots for it (the HBM budget is the context's slot capacity: every
field is allocated for all slots, so a slot costs the same bytes whatever case holds it).  All
admitted cases advance together: one ffvhc_step call per loop iteration runs every stage of every
active case in one launch per stage, each case with its own Poisson convergence and step counter.
A case leaves when it has run its steps; its slots are reused by the next admission.  Waiting
cases are tried in arrival order at every step; a later, smaller case may start while an earlier
one still waits for a large enough run of slots (backfill).

Time-to-solution of a case = completion time - arrival time (wall clock).  The per-case results
do not depend on the mix: every kernel works per block and every reduction per case in a fixed
order, so a case's fields are bit-identical alone and in any mix.
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
    steps: int                # time steps to run
    init: Callable            # init(ctx, case_id, fields) uploads the initial condition
    geometry: Optional[tuple] = None
    result: Optional[Callable] = None  # result(ctx, case_id, fields) before the case is removed
    # filled by the loop
    admitted: float = -1.0
    finished: float = -1.0
    case_id: int = -1
    done_steps: int = 0
    poisson_iters: List[int] = field(default_factory=list)


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


class Server:
    """Admit, interleave and retire cases on one context (continuous batching)."""

    def __init__(self, ctx, fields, step_kwargs, clock=time.perf_counter):
        self.ctx = ctx
        self.F = fields
        self.step_kwargs = dict(step_kwargs)
        self.clock = clock
        self.waiting: List[Request] = []
        self.active: List[Request] = []
        self.finished: List[Request] = []
        self.steps_run = 0
        self.cell_steps = 0

    def _admit(self, now):
        still = []
        for r in self.waiting:
            try:
                cid = self.ctx.add_case(r.tree, r.params, geometry=r.geometry)
            except Exception as e:  # status -3: no contiguous room yet, keep waiting (FIFO order kept)
                if getattr(e, "code", None) != -3:
                    raise
                still.append(r)
                continue
            r.case_id = cid
            r.admitted = now
            self.ctx.select([cid])
            r.init(self.ctx, cid, self.F)
            self.active.append(r)
        self.waiting = still

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
            now = self.clock() - t0
            keep = []
            for r, inf in zip(self.active, info):
                r.done_steps += 1
                r.poisson_iters.append(int(inf.countP))
                self.cell_steps += r.tree.nblk * 16 ** 3
                if r.done_steps >= r.steps:
                    r.finished = now
                    if r.result is not None:
                        self.ctx.select([r.case_id])
                        r.result(self.ctx, r.case_id, self.F)
                    self.ctx.remove_case(r.case_id)
                    self.finished.append(r)
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
               "mcell_steps_s": self.cell_steps / max(self.wall, 1e-12) / 1e6}
        if tts.size:
            out.update(p50_s=float(np.percentile(tts, 50)), p99_s=float(np.percentile(tts, 99)),
                       cases_per_hour=len(self.finished) / max(self.wall, 1e-12) * 3600.0)
        return out
