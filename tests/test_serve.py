"""Host logic of the continuous-batching loop (admission by slot capacity, FIFO waiting,
interleaved steps, retirement, latency statistics), driven by a recording stand-in context."""
import numpy as np
import pytest

from serve import Request, Server, bytes_per_block


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
        return [Info(7) for _ in self.sel], np.zeros(7), 0


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
