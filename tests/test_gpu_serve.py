"""Served stream on the GPU: a case's fields after its steps are bit-identical whether it runs
alone or is admitted into a running mix of other cases of different sizes (continuous batching
with per-case reductions)."""
import numpy as np
import pytest

import ffvhc
from serve import Request, Server, uniform_flow_init

pytestmark = pytest.mark.gpu

BC = {ffvhc.BC_UX: [(0, 1.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)],
      ffvhc.BC_P: [(1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)]}


class StepClock:
    """Deterministic clock: advances a fixed amount per call."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.25
        return self.t


def tree(kind):
    if kind == "a":
        return ffvhc.Tree(root=(2, 2, 2), root_length=1.0, regions=[dict(kind="ball", a=(1.0, 1.0, 1.0), r=0.3,
                                                                          level=1)])
    if kind == "b":
        return ffvhc.Tree(root=(4, 2, 2), root_length=1.0)
    return ffvhc.Tree(root=(3, 3, 2), root_length=1.0, regions=[dict(kind="ball", a=(1.5, 1.5, 1.0), r=0.4,
                                                                      level=1)])


def serve(plan, step_kwargs, cap=2048):
    ctx = ffvhc.Context(max_blocks=cap)
    F = ctx.step_fields()
    out = {}

    def grab(name):
        def result(ctx_, cid, F_):
            out[name] = [ctx_.download(f, cid, out=np.zeros((ctx_.cases[cid].nblk, 20, 20, 20)))
                         for f in (F_.ux, F_.uy, F_.uz, F_.p, F_.lapp)]
        return result

    reqs = [Request(name=n, arrival=a, tree=tree(k), params=ffvhc.case_params(mu=0.01, dt=0.005, bc=BC),
                    steps=s, init=uniform_flow_init(seed), result=grab(n))
            for n, k, a, s, seed in plan]
    srv = Server(ctx, F, step_kwargs, clock=StepClock())
    srv.run(reqs)
    rep = srv.report()
    ctx.close()
    return out, rep, reqs


@pytest.mark.parametrize("solver", [0, 3, "m15", "m19", "m26"])
def test_case_alone_equals_case_in_mix(solver):
    if solver in ("m15", "m19", "m26"):  # FP64 BiCGSTAB on the FP32 matrix, SOR sweeps (+ coarse correction)
        kw = dict(solver=3, omega=1.6, count_max=300, eps=1e-8, mixed=True, mixed_method=int(solver[1:]))
    else:
        kw = dict(solver=solver, omega=1.2 if solver == 0 else 1.0, count_max=300, eps=1e-6)
    alone, _, _ = serve([("A", "a", 0.0, 4, 1)], kw)
    mix, rep, reqs = serve([("B", "b", 0.0, 6, 2), ("C", "c", 0.0, 3, 3), ("A", "a", 1.0, 4, 1)], kw)
    # A was admitted into a running mix
    a = [r for r in reqs if r.name == "A"][0]
    assert a.admitted > 0.0
    assert rep["cases"] == 3
    for x, y in zip(alone["A"], mix["A"]):
        assert np.array_equal(x, y)


def test_drag_stopped_case_alone_equals_case_in_mix():
    """A flow past a sphere under the drag rule (here: stop once Cd changes by < 0.2 over 5 steps,
    at most 40 steps) stops at the same step with the same drag history and fields, bit for bit,
    alone and admitted into a running mix of other cases."""
    import flows
    from serve import DragStop

    def sphere_req(name, arrival, target, levels, re, seed, out):
        tree, center = flows.served_case(target, levels)
        geo = tree.geometry_sphere(center=center, radius=0.5, cut_id=1)[:3]

        def grab(ctx_, cid, F_):
            out[name] = [ctx_.download(f, cid, out=np.zeros((ctx_.cases[cid].nblk, 20, 20, 20)))
                         for f in (F_.ux, F_.uy, F_.uz, F_.p)]
        return Request(name=name, arrival=arrival, tree=tree, params=flows.flow_params(re, flows.cfl_dt(tree, 0.3)),
                       steps=40, init=uniform_flow_init(seed, noise=0.0), geometry=geo, result=grab,
                       drag_stop=DragStop(window=5, tol=0.2))

    kw = dict(solver=3, omega=1.6, count_max=2000, eps=1e-8, mixed=True, mixed_method=26, implicit_u=True,
              alpha_u=1.0, omega_u=1.0, count_max_u=1000, eps_u=1e-8)

    def run(plan):
        ctx = ffvhc.Context(max_blocks=4096)
        F = ctx.step_fields()
        out = {}
        reqs = [sphere_req(*p, out=out) for p in plan]
        srv = Server(ctx, F, kw, clock=StepClock())
        srv.run(reqs)
        ctx.close()
        return out, reqs

    alone, ra = run([("A", 0.0, 100, 2, 120.0, 1)])
    mix, rm = run([("B", 0.0, 300, 1, 80.0, 2), ("C", 0.0, 150, 3, 250.0, 3), ("A", 1.0, 100, 2, 120.0, 1)])
    a1 = [r for r in ra if r.name == "A"][0]
    a2 = [r for r in rm if r.name == "A"][0]
    assert a2.admitted > 0.0                      # admitted into a running mix
    print("case A: %d steps (%s), Cd %s" % (a1.done_steps, a1.stop_reason, a1.drag[-3:]))
    assert len(a1.drag) == a1.done_steps and all(np.isfinite(a1.drag))
    assert (a1.done_steps, a1.stop_reason) == (a2.done_steps, a2.stop_reason)
    assert a1.drag == a2.drag
    for x, y in zip(alone["A"], mix["A"]):
        assert np.array_equal(x, y)
