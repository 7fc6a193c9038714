"""Host logic of the continuous-batching loop (admission by slot capacity, FIFO waiting,
interleaved steps, retirement, latency statistics), driven by a recording stand-in context."""
import numpy as np
import pytest

from serve import DragStop, Request, Server, bytes_per_block, coarse_bytes, drag_settled


class CapacityError(RuntimeError):
    code = -3


class FakeTree:
    def __init__(self, nblk):
        self.nblk = nblk


class Info:
    def __init__(self, n):
        self.countP = n


class FakeCtx:
    """Slot allocator with first-fit contiguous ranges, like ffvhc_case_add."""

    def __init__(self, cap):
        self.cap = cap
        self.steps_of = {}
        self.used = {}
        self.next_id = 0
        self.sel = None
        self.log = []

    def add_case(self, tree, params, geometry=None):
        taken = sorted(self.used.values())
        begin, pos = None, 0
        for b, e in taken + [(self.cap, self.cap)]:
            if b - pos >= tree.nblk:
                begin = pos
                break
            pos = max(pos, e)
        if begin is None:
            raise CapacityError("not enough free block slots")
        cid = self.next_id
        self.next_id += 1
        self.used[cid] = (begin, begin + tree.nblk)
        return cid

    def remove_case(self, cid):
        del self.used[cid]

    def select(self, ids):
        self.sel = None if ids is None else list(ids)

    def step(self, F, ncases=None, **kw):
        assert ncases == len(self.sel)
        self.log.append(list(self.sel))
        for c in self.sel:
            self.steps_of[c] = self.steps_of.get(c, 0) + 1
        return [Info(7) for _ in self.sel], np.zeros(7), 0

    # forces: x force 1 + 0.5^n after n steps of the case (pressure and viscous halves), so the
    # drag coefficient settles geometrically
    steps_of = None

    def force_p_noalloc(self, p, cut_id, ncases=None):
        assert ncases == len(self.sel)
        return np.array([[0.5 * (1.0 + 0.5 ** self.steps_of[c]), 0.0, 0.0] for c in self.sel])

    def force_v(self, ux, uy, uz, cut_id, ncases=None):
        return self.force_p_noalloc(None, cut_id, ncases)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def reqs(sizes, arrivals, steps):
    return [Request(name="c%d" % i, arrival=a, tree=FakeTree(n), params=None, steps=s,
                    init=lambda ctx, cid, F: None) for i, (n, a, s) in enumerate(zip(sizes, arrivals, steps))]


def test_all_cases_finish_with_their_step_counts():
    ctx = FakeCtx(cap=100)
    srv = Server(ctx, fields=None, step_kwargs={}, clock=Clock())
    rs = reqs([40, 30, 50, 20], [0.0, 0.0, 0.0, 0.0], [3, 5, 2, 4])
    done = srv.run(rs)
    assert sorted(r.name for r in done) == ["c0", "c1", "c2", "c3"]
    for r in done:
        assert r.done_steps == r.steps and len(r.poisson_iters) == r.steps
    assert not ctx.used  # every slot returned


def test_admission_waits_for_capacity_and_backfills():
    ctx = FakeCtx(cap=100)
    srv = Server(ctx, fields=None, step_kwargs={}, clock=Clock())
    rs = reqs([60, 60, 30], [0.0, 0.0, 0.0], [2, 2, 2])
    srv.run(rs)
    c0, c1, c2 = rs
    # c1 does not fit beside c0 and waits; the smaller c2 fills the room left (backfill)
    assert c0.admitted == c2.admitted < c1.admitted
    assert c1.admitted >= c0.finished
    assert all(len(s) <= 2 for s in ctx.log)


def test_interleaving_runs_cases_in_one_step_call():
    ctx = FakeCtx(cap=1000)
    srv = Server(ctx, fields=None, step_kwargs={}, clock=Clock())
    srv.run(reqs([10, 20, 30], [0.0, 0.0, 0.0], [4, 4, 4]))
    assert ctx.log[0] == [0, 1, 2]
    assert len(ctx.log) == 4


def test_report_percentiles():
    ctx = FakeCtx(cap=1000)
    srv = Server(ctx, fields=None, step_kwargs={}, clock=Clock())
    srv.run(reqs([10] * 10, [0.0] * 10, list(range(1, 11))))
    rep = srv.report()
    assert rep["cases"] == 10
    assert rep["p50_s"] <= rep["p99_s"]
    assert rep["cases_per_hour"] > 0


def test_case_that_can_never_fit_is_an_error():
    ctx = FakeCtx(cap=10)
    srv = Server(ctx, fields=None, step_kwargs={}, clock=Clock())
    with pytest.raises(RuntimeError):
        srv.run(reqs([20], [0.0], [1]))


def test_bytes_per_block():
    assert bytes_per_block(1) == 8 * (4096 + 1536)
    assert bytes_per_block(0, 2) == 8 * (4096 + 1536)


class Params:
    rho = 2.0


def test_drag_rule_stops_a_case_once_settled():
    """Cd = F_x / (rho u0^2 A / 2); with F_x = 1 + 0.5^n the change over a window of 4 steps falls
    below 1e-3 once 0.5^(n-4) < 1e-3 (n = 14)."""
    ctx = FakeCtx(cap=100)
    srv = Server(ctx, fields=type("F", (), {"p": 0, "ux": 1, "uy": 2, "uz": 3})(), step_kwargs={}, clock=Clock())
    rs = [Request(name="s", arrival=0.0, tree=FakeTree(10), params=Params(), steps=1000, init=lambda c, i, F: None,
                  drag_stop=DragStop(window=4, tol=1e-3, area=0.5, u0=1.0)),
          Request(name="f", arrival=0.0, tree=FakeTree(10), params=Params(), steps=3, init=lambda c, i, F: None)]
    srv.run(rs)
    s_, f_ = rs
    assert f_.stop_reason == "step limit" and f_.done_steps == 3 and not f_.drag
    assert s_.stop_reason == "drag settled"
    cd = [(1.0 + 0.5 ** n) / (0.5 * 2.0 * 1.0 * 0.5) for n in range(1, s_.done_steps + 1)]
    assert np.allclose(s_.drag, cd)
    want = next(n for n in range(1, 100) if drag_settled(cd[:n], 4, 1e-3))
    assert s_.done_steps == want
    rep = srv.report()
    assert rep["stopped_by_drag"] == 1 and rep["stopped_by_step_limit"] == 1


def test_pool_budget_delays_admission():
    """Two cases whose coarse operators do not fit the pool together run one after the other, even
    though their slots fit."""
    ctx = FakeCtx(cap=1000)
    pool = coarse_bytes(300) + coarse_bytes(100)
    srv = Server(ctx, fields=None, step_kwargs={}, clock=Clock(), pool_bytes=pool)
    rs = reqs([300, 250, 100], [0.0, 0.0, 0.0], [2, 2, 2])
    srv.run(rs)
    a, b, c = rs
    assert a.admitted == c.admitted < b.admitted  # c backfills beside a; b waits for the pool
    assert b.admitted >= a.finished
    rep = srv.report()
    assert rep["peak_pool_bytes"] <= pool
    assert coarse_bytes(4) == 8 * 16 + 4 * 16 and coarse_bytes(5) == 8 * 25 + 4 * 5 * 8
