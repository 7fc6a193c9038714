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
