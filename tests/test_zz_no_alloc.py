"""The allocation-free pressure-force entry point (ffvhc_force_p_noalloc) against ffvhc_force_p.

Both run the same kernels in the same order; the allocation-free one keeps its block partials and
case triples in a fixed per-device scratch array, which keeps the last call's contents.  So two
calls in a row, the second on a smaller input (fewer blocks, fewer cases), must each give exactly
the per-call-allocating entry point's triples."""
import numpy as np
import pytest

import ffvhc
from helpers import OracleGrid, geometry

pytestmark = pytest.mark.gpu

BC = {ffvhc.BC_P: [(1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)]}


def _case(ctx, tree, seed):
    og = OracleGrid(tree)
    rng = np.random.default_rng(seed)
    geo = geometry(og, rng)
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params, geometry=(geo[0], geo[1], geo[2]))
    p = rng.standard_normal(og.shape)
    og.ghost_fill(p, 1, [params.bct[ffvhc.BC_P][f] for f in range(6)],
                  [params.bcv[ffvhc.BC_P][f] for f in range(6)])
    return cid, p


def test_force_p_noalloc_equals_force_p_on_two_calls():
    big = ffvhc.Tree(root=(3, 2, 2), root_length=1.0,
                     regions=[dict(kind="ball", a=(1.5, 1.0, 1.0), r=0.4, level=1)])
    small = ffvhc.Tree(root=(2, 1, 1), root_length=1.0)
    assert small.nblk < big.nblk
    ctx = ffvhc.Context(max_blocks=big.nblk + small.nblk)
    pf = ctx.field(8)
    try:
        cb, pb = _case(ctx, big, 31)
        cs, ps = _case(ctx, small, 32)
        ctx.upload(pf, cb, pb)
        ctx.upload(pf, cs, ps)
        # first call: the larger case and the small one together, then the smaller input alone
        for sel in ([cb, cs], [cs]):
            ctx.select(sel)
            for target in (1, 2):
                want = ctx.force_p(pf, target, ncases=len(sel))
                got = ctx.force_p_noalloc(pf, target, ncases=len(sel))
                assert np.array_equal(got, want), (sel, target, got, want)
                assert np.any(want != 0.0)
    finally:
        ctx.select(None)
        ctx.free(pf)
        ctx.close()
