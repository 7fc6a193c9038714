"""Selection logic of the schedule autotuner with a stand-in context (timings are simulated)."""
import numpy as np
import pytest

import autotune


class FakeCtx:
    def __init__(self, behaviour):
        self.behaviour = behaviour  # (method, ieps) -> (seconds, converged)
        self.now = 0.0

    def fill(self, x, v):
        pass

    def sync(self):
        pass

    def solve_mixed(self, method, x, A, b, group, omega, cap, eps, ieps, ncases=1):
        sec, ok = self.behaviour[(method, ieps)]
        self.now += sec
        res = np.array([eps * 0.5 if ok else 1.0])
        return np.array([100]), res, np.array([3]), 0 if ok else 1


def test_picks_fastest_converged(monkeypatch):
    beh = {(3, 0.1): (0.5, True), (3, 0.01): (0.2, False), (1, 0.01): (0.3, True)}
    ctx = FakeCtx(beh)

    def fake_perf():
        return ctx.now

    monkeypatch.setattr(autotune.time, "perf_counter", fake_perf)
    cands = ((3, 0.1, 10, 1.0), (3, 0.01, 10, 1.0), (1, 0.01, 10, 1.0))
    best, res = autotune.tune_mixed(ctx, 0, [0] * 7, 0, 3, 1e-10, candidates=cands)
    assert best == (1, 0.01, 10, 1.0)
    assert res[1][1] == float("inf")


def test_no_candidate_converges(monkeypatch):
    ctx = FakeCtx({(3, 0.1): (0.5, False)})
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: ctx.now)
    with pytest.raises(RuntimeError):
        autotune.tune_mixed(ctx, 0, [0] * 7, 0, 3, 1e-10, candidates=((3, 0.1, 10, 1.0),))


class FakeCtxPerRhs(FakeCtx):
    def solve_mixed(self, method, x, A, b, group, omega, cap, eps, ieps, ncases=1):
        sec, ok = self.behaviour[(method, b)]
        self.now += sec
        return np.array([40 + b]), np.array([eps * 0.5]), np.array([1]), 0


def test_several_right_hand_sides_are_summed(monkeypatch):
    """A schedule that wins on one right-hand side but loses on the others is not picked: the
    time is the sum over the right-hand sides, the iterations the largest count."""
    beh = {(26, 1): (0.010, True), (26, 2): (0.010, True), (26, 3): (0.010, True),
           (27, 1): (0.009, True), (27, 2): (0.012, True), (27, 3): (0.011, True)}
    ctx = FakeCtxPerRhs(beh)
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: ctx.now)
    cands = ((27, 0.5, 20000, 1.6), (26, 0.5, 20000, 1.6))
    best, res = autotune.tune_mixed(ctx, 0, [0] * 7, [1, 2, 3], 3, 1e-10, candidates=cands, repeats=2)
    assert best == (26, 0.5, 20000, 1.6)
    assert res[0][1] == pytest.approx(0.032) and res[1][1] == pytest.approx(0.030)
    assert res[0][2] == 43
    best1, _ = autotune.tune_mixed(ctx, 0, [0] * 7, 1, 3, 1e-10, candidates=cands)
    assert best1 == (27, 0.5, 20000, 1.6)
