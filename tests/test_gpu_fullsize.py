"""The benchmark's own system at full size (1,128 blocks of 16^3, two levels), checked through
properties that hold at any size: every solver reaches the FP64 relative residual it reports
(recomputed by ffvhc_residual), the fast solvers' answers equal the FP64 BiCGSTAB answer to the
solver tolerance, and the solves are homogeneous (b -> 2b gives exactly 2x, bit for bit: scaling
by two is exact in binary floating point and every step of the solvers is linear in b)."""
import os
import sys

import numpy as np
import pytest

import ffvhc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.fixture(scope="module")
def c2():
    tree = bench.build_c2()
    A, b, bct = bench.poisson_system(tree)
    ctx, cid, f64, f32 = bench.setup(tree, A, b, bct)
    yield ctx, cid, f64, tree, b
    ctx.close()


def solve(ctx, f64, method):
    ctx.fill(f64["x"], 0.0)
    if method is None:
        cnt, res, rc = ctx.bicgstab(f64["x"], f64["A"], f64["b"], ffvhc.BC_P, 1.0, 1, 20000, 1e-10, ncases=1)
    elif method == 29:  # FP32 inner solves to 1e-5 per correction, at most 200 iterations each
        cnt, res, _o, rc = ctx.solve_mixed(method, f64["x"], f64["A"], f64["b"], ffvhc.BC_P, 1.6, 200, 1e-10,
                                           1e-5, ncases=1)
    else:
        cnt, res, _o, rc = ctx.solve_mixed(method, f64["x"], f64["A"], f64["b"], ffvhc.BC_P, 1.6, 20000, 1e-10,
                                           0.5, ncases=1)
    assert rc == 0
    return int(cnt[0]), float(res[0])


def test_fullsize_solvers_reach_and_agree(c2):
    ctx, cid, f64, tree, b = c2
    shp = (tree.nblk, 20, 20, 20)
    inner = (slice(None),) + (slice(2, 18),) * 3
    sols = {}
    for method in (None, 15, 19, 25, 26, 29):
        cnt, res = solve(ctx, f64, method)
        true = ctx.residual(f64["x"], f64["A"], f64["b"], ncases=1)[0]
        assert res <= 1e-10 and true <= 1e-10 * 1.0001, (method, cnt, res, true)
        sols[method] = ctx.download(f64["x"], cid, out=np.zeros(shp))[inner]
    scale = np.max(np.abs(sols[None]))
    for m in (15, 19, 25, 26, 29):
        assert np.max(np.abs(sols[m] - sols[None])) <= 1e-7 * scale


@pytest.mark.parametrize("method", [None, 19, 25, 26, 29])
def test_fullsize_solve_is_homogeneous(c2, method):
    ctx, cid, f64, tree, b = c2
    shp = (tree.nblk, 20, 20, 20)
    inner = (slice(None),) + (slice(2, 18),) * 3
    c1, _ = solve(ctx, f64, method)
    x1 = ctx.download(f64["x"], cid, out=np.zeros(shp))[inner]
    ctx.upload(f64["b"], cid, 2.0 * b)
    try:
        c2_, _ = solve(ctx, f64, method)
        x2 = ctx.download(f64["x"], cid, out=np.zeros(shp))[inner]
    finally:
        ctx.upload(f64["b"], cid, b)
    assert c1 == c2_
    assert np.array_equal(x2, 2.0 * x1)


def test_fullsize_coarse_inverse_gemm_and_vector_paths(c2):
    """The coarse operator of the 1,128-block system inverted both ways: 256-wide panels with FP64
    GEMM products (five panels, the last one partial) and 32-wide panels on the vector units.  Both
    invert A_c (A_c^-1 A_c = I to 1e-9) and agree with each other to rounding."""
    ctx, cid, f64, tree, b = c2
    n = tree.nblk
    out = {}
    try:
        for gemm_min in (1, 1 << 30):
            ctx.set_coarse_gemm_min(gemm_min)
            ac, ainv = ctx.coarse_operator(f64["A"], cid)
            assert ac.shape == (n, n)
            err = np.max(np.abs(ainv @ ac - np.eye(n)))
            assert err <= 1e-9, (gemm_min, err)
            out[gemm_min] = ainv
    finally:
        ctx.set_coarse_gemm_min(2048)
    scale = np.max(np.abs(out[1]))
    assert np.max(np.abs(out[1] - out[1 << 30])) <= 1e-10 * scale


def test_large_case_fused_finalisation_bit_identical():
    """A case of 4,352 blocks (flat 17 x 16 x 16 roots), above the 4,096 blocks whose coarse
    right-hand side the coarse application stages in LDS: every wave then forms the coarse
    right-hand side and the previous iteration's end itself (17 strided rounds of the per-wave
    sum).  Method 26 with the finalisation fused and in its own launch: the same count and every
    bit of x, and the FP64 residual recomputed from the answer meets the tolerance."""
    tree = ffvhc.Tree(root=(17, 16, 16), root_length=1.0)
    assert tree.nblk == 4352
    A, b, bct = bench.poisson_system(tree)
    ctx, cid, f64, f32 = bench.setup(tree, A, b, bct)
    del A
    try:
        shp = (tree.nblk, 20, 20, 20)
        got = []
        for fuse in (True, False):
            ctx.set_fuse_fin(fuse)
            cnt, res = solve(ctx, f64, 26)
            true = ctx.residual(f64["x"], f64["A"], f64["b"], ncases=1)[0]
            assert res <= 1e-10 and true <= 1e-10 * 1.0001, (fuse, cnt, res, true)
            got.append((cnt, res, ctx.download(f64["x"], cid, out=np.zeros(shp))))
        assert got[0][0] == got[1][0] and got[0][1] == got[1][1]
        assert np.array_equal(got[0][2], got[1][2])
    finally:
        ctx.set_fuse_fin(True)
        ctx.close()
