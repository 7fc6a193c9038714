"""GPU parity of the ghost fill and the fused RBGS iteration against the oracle (FP64).

Bit-exact: the device kernels evaluate the same expressions in the same order with no FMA
contraction; only reduction order differs (residual norms, checked to 1e-12 relative).
Compared region: interior cells and the first ghost layer of each face (what the device keeps).
"""
import ctypes

import numpy as np
import pytest

import ffvhc
from oracle import oracle as O
from helpers import G, N, OracleGrid, cell_centres, interior, live_mask

pytestmark = pytest.mark.gpu

LIVE = live_mask()


def trees():
    return {
        "flat": ffvhc.Tree(root=(3, 2, 2), root_length=1.0),
        "two_level": ffvhc.Tree(root=(3, 3, 2), root_length=1.0,
                                regions=[dict(kind="ball", a=(1.4, 1.6, 1.0), r=0.45, level=1)]),
        "three_level": ffvhc.Tree(root=(2, 2, 2), root_length=1.0,
                                  regions=[dict(kind="ball", a=(0.9, 1.1, 1.05), r=0.2, level=2)]),
    }


BC = {ffvhc.BC_UX: [(0, 1.0), (1, 0.0), (1, 0.0), (1, 0.25), (0, -0.5), (1, 0.0)],
      ffvhc.BC_P: [(1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)]}


def assert_live_equal(got, want, what):
    bad = np.flatnonzero((got != want)[:, LIVE])
    assert bad.size == 0, "%s: %d live cells differ, max |d| %.3e" % (
        what, bad.size, np.max(np.abs((got - want)[:, LIVE])))


@pytest.fixture(scope="module")
def ctx():
    c = ffvhc.Context(max_blocks=4096)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["flat", "two_level", "three_level"])
@pytest.mark.parametrize("updater,group", [(1, ffvhc.BC_P), (5, ffvhc.BC_UX)])
def test_ghost_fill_bit_exact(ctx, name, updater, group):
    tree = trees()[name]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(7)
        f = rng.uniform(-1, 1, og.shape)
        fld = ctx.field(8)
        ctx.upload(fld, cid, f)
        ctx.ghost_fill(fld, group, updater)
        got = ctx.download(fld, cid, out=f.copy())
        bct = [params.bct[group][q] for q in range(6)]
        bcv = [params.bcv[group][q] for q in range(6)]
        og.ghost_fill(f, updater, bct, bcv)
        assert_live_equal(got, f, "ghost fill %s u%d" % (name, updater))
        ctx.free(fld)
    finally:
        ctx.remove_case(cid)


def make_system(og, rng, bct, bcv):
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = rng.uniform(0.6, 1.4, og.shape)
    A[0][:] = -(sum(A[q] for q in range(1, 7))) - rng.uniform(0.0, 0.2, og.shape)
    b = rng.uniform(-1, 1, og.shape)
    og.L.orc_matrix_bc(og.ref(), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv))
    return A, b


def run_oracle_rbgs(og, x, A, b, bct, bcv, omega, cmax, eps):
    cnt, res = ctypes.c_int(), ctypes.c_double()
    og.L.orc_rbgs(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv), ctypes.c_double(omega), cmax,
                  ctypes.c_double(eps), ctypes.byref(cnt), ctypes.byref(res))
    return cnt.value, res.value


def upload_system(ctx, cid, A, b, x):
    fa = [ctx.field(8) for _ in range(7)]
    for q in range(7):
        ctx.upload(fa[q], cid, A[q])
    fb, fx = ctx.field(8), ctx.field(8)
    ctx.upload(fb, cid, b)
    ctx.upload(fx, cid, x)
    return fa, fb, fx


@pytest.mark.parametrize("name", ["flat", "two_level", "three_level"])
def test_rbgs_fixed_count_bit_exact(ctx, name):
    tree = trees()[name]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(3)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        x = rng.uniform(-1, 1, og.shape)
        og.ghost_fill(x, 1, bct, bcv)
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.rbgs(fx, fa, fb, ffvhc.BC_P, 1.2, 9, 0.0, ncases=1)
        got = ctx.download(fx, cid, out=x.copy())
        c2, r2 = run_oracle_rbgs(og, x, A, b, bct, bcv, 1.2, 9, 0.0)
        assert rc == 0 and cnt[0] == c2 == 10
        assert abs(res[0] - r2) <= 1e-12 * abs(r2)
        assert_live_equal(got, x, "rbgs %s" % name)
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


@pytest.mark.parametrize("name", ["flat", "three_level"])
def test_jacobi_fixed_count_bit_exact(ctx, name):
    """FFVILS::Jacobi: every update is local, so a fixed sweep count is bit-exact."""
    tree = trees()[name]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(11)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        x = rng.uniform(-1, 1, og.shape)
        og.ghost_fill(x, 1, bct, bcv)
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.jacobi(fx, fa, fb, ffvhc.BC_P, 0.8, 7, 0.0, ncases=1)
        got = ctx.download(fx, cid, out=x.copy())
        c2, r2 = ctypes.c_int(), ctypes.c_double()
        og.L.orc_jacobi(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.ptr(og.zeros()), O.cint6(bct), O.cdbl6(bcv),
                        ctypes.c_double(0.8), 7, ctypes.c_double(0.0), ctypes.byref(c2), ctypes.byref(r2))
        assert rc == 0 and cnt[0] == c2.value == 8
        assert abs(res[0] - r2.value) <= 1e-12 * abs(r2.value)
        assert_live_equal(got, x, "jacobi %s" % name)
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


def test_rbgs_fixed_average_matches(ctx):
    """FFVILS::RBGS_WithFixedAverage on an all-Neumann system: the mean is removed every
    iteration; the mean is a global sum, so agreement is to 1e-12 (sum order differs)."""
    tree = trees()["two_level"]
    bc = {ffvhc.BC_P: [(1, 0.0)] * 6}
    params = ffvhc.case_params(bc=bc)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(12)
        bct, bcv = [1] * 6, [0.0] * 6
        A, b = make_system(og, rng, bct, bcv)
        x = og.zeros()
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.rbgs_fixed_average(fx, fa, fb, ffvhc.BC_P, 1.2, 12, 0.0, ncases=1)
        got = ctx.download(fx, cid, out=x.copy())
        c2, r2 = ctypes.c_int(), ctypes.c_double()
        og.L.orc_rbgs_fixed_average(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv),
                                    og.nb * N ** 3, ctypes.c_double(1.2), 12, ctypes.c_double(0.0),
                                    ctypes.byref(c2), ctypes.byref(r2))
        assert rc == 0 and cnt[0] == c2.value == 13
        assert abs(res[0] - r2.value) <= 1e-10 * abs(r2.value)
        scale = np.max(np.abs(x[:, LIVE]))
        assert np.max(np.abs(got - x)[:, LIVE]) <= 1e-12 * scale
        assert abs(np.mean(interior(got))) <= 1e-12 * scale
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


def test_rbgs_converges_like_oracle(ctx):
    tree = trees()["two_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(4)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        x = og.zeros()
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.rbgs(fx, fa, fb, ffvhc.BC_P, 1.2, 500, 1e-6, ncases=1)
        got = ctx.download(fx, cid, out=x.copy())
        c2, r2 = run_oracle_rbgs(og, x, A, b, bct, bcv, 1.2, 500, 1e-6)
        assert cnt[0] == c2 and c2 < 500
        assert res[0] <= 1e-6
        assert np.allclose(got[:, LIVE], x[:, LIVE], rtol=0, atol=1e-12 * np.max(np.abs(x)))
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


def test_two_cases_in_one_launch_equal_alone(ctx):
    """Batched solve of two cases of different sizes: each result is bit-identical to its solo run."""
    ta, tb_ = trees()["two_level"], trees()["flat"]
    params = ffvhc.case_params(bc=BC)
    bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
    bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
    results = {}
    for mode in ("alone", "mix"):
        ids = []
        fields = []
        for k, tr in enumerate((ta, tb_)):
            if mode == "alone" and k == 1:
                continue
            ids.append(ctx.add_case(tr, params))
        if mode == "alone":
            ids.append(None)
        systems = []
        for k, tr in enumerate((ta, tb_)):
            og = OracleGrid(tr)
            rng = np.random.default_rng(20 + k)
            A, b = make_system(og, rng, bct, bcv)
            systems.append((og, A, b))
        fa = [ctx.field(8) for _ in range(7)]
        fb, fx = ctx.field(8), ctx.field(8)
        fields = fa + [fb, fx]
        outs = []
        if mode == "mix":
            for k in range(2):
                og, A, b = systems[k]
                for q in range(7):
                    ctx.upload(fa[q], ids[k], A[q])
                ctx.upload(fb, ids[k], b)
                ctx.upload(fx, ids[k], og.zeros())
            cnt, res, rc = ctx.rbgs(fx, fa, fb, ffvhc.BC_P, 1.1, 300, 1e-5, ncases=2)
            outs = [(ctx.download(fx, ids[k]), cnt[k], res[k]) for k in range(2)]
            for i in ids:
                ctx.remove_case(i)
        else:
            ctx.remove_case(ids[0])
            for k in range(2):
                i = ctx.add_case((ta, tb_)[k], params)
                og, A, b = systems[k]
                for q in range(7):
                    ctx.upload(fa[q], i, A[q])
                ctx.upload(fb, i, b)
                ctx.upload(fx, i, og.zeros())
                cnt, res, rc = ctx.rbgs(fx, fa, fb, ffvhc.BC_P, 1.1, 300, 1e-5, ncases=1)
                outs.append((ctx.download(fx, i), cnt[0], res[0]))
                ctx.remove_case(i)
        results[mode] = outs
        for f in fields:
            ctx.free(f)
    for k in range(2):
        xa, ca, ra = results["alone"][k]
        xm, cm, rm = results["mix"][k]
        assert ca == cm and ra == rm
        assert np.array_equal(xa[:, LIVE], xm[:, LIVE])


def run_oracle_cg(og, x, A, b, bct, bcv, omega, cmax, eps):
    cnt, res = ctypes.c_int(), ctypes.c_double()
    scratch = [og.zeros() for _ in range(5)]
    og.L.orc_cg(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), *[O.ptr(s) for s in scratch], O.cint6(bct), O.cdbl6(bcv),
                ctypes.c_double(omega), 1, cmax, ctypes.c_double(eps), ctypes.byref(cnt), ctypes.byref(res))
    return cnt.value, res.value


def spd_system(og, rng, bct, bcv):
    """Symmetric 7-point system (uniform bands per block), outer BC folded; single level keeps it SPD."""
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = 1.0
    A[0][:] = -6.0 - rng.uniform(0.0, 0.05, og.shape)
    b = rng.uniform(-1, 1, og.shape)
    og.L.orc_matrix_bc(og.ref(), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv))
    return A, b


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_cg_fixed_count_matches_oracle(ctx, name):
    tree = trees()[name]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(8)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = spd_system(og, rng, bct, bcv)
        x = og.zeros()
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.cg(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 15, 0.0, ncases=1)
        got = ctx.download(fx, cid, out=x.copy())
        c2, r2 = run_oracle_cg(og, x, A, b, bct, bcv, 1.0, 15, 0.0)
        assert rc == 0 and cnt[0] == c2 == 16
        assert abs(res[0] - r2) <= 1e-9 * abs(r2)
        scale = np.max(np.abs(x[:, LIVE]))
        assert np.max(np.abs(got[:, LIVE] - x[:, LIVE])) <= 1e-12 * scale
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


def true_residual(og, A, x, b):
    rr = og.L.orc_grid_r2(og.ref(), O.parr(A), O.ptr(x), O.ptr(b))
    bb = og.L.orc_grid_dot(og.ref(), O.ptr(b), O.ptr(b))
    return np.sqrt(rr / bb)


@pytest.mark.parametrize("method", [0, 1, 3, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 17, 19, 21, 23, 25, 26, 27, 29])
def test_mixed_precision_reaches_fp64_residual(ctx, method):
    tree = trees()["two_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(9)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = spd_system(og, rng, bct, bcv) if method == 1 else make_system(og, rng, bct, bcv)
        x = og.zeros()
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        eps = 1e-10
        # methods 0/1: inner relative reduction 1e-4; method 3: reliable update every 1e-2; methods 5-9
        # (FP32-stored matrix / preconditioner) run on the random coefficients here, which FP32 does not
        # hold exactly, so their acceptance test on the FP64 residual is exercised
        cnt, res, outer, rc = ctx.solve_mixed(method, fx, fa, fb, ffvhc.BC_P, 1.0 if method else 1.2, 100000, eps,
                                              1e-2 if method == 3 else 1e-4, ncases=1)
        assert rc == 0 and res[0] <= eps and outer[0] >= (2 if method < 5 or method == 29 else 0)
        got = ctx.download(fx, cid, out=x.copy())
        # the FP64 residual the device reports is the true one (computed by the oracle on the result)
        tr = true_residual(og, A, got, b)
        assert tr <= eps * 1.0001 and abs(tr - res[0]) <= 1e-3 * res[0] + 1e-16
        # and it is the FP64 solution: compare with the FP64 solver at the same tolerance
        ctx.upload(fx, cid, og.zeros())
        if method == 1:
            c64, r64, _ = ctx.cg(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 100000, eps, ncases=1)
        elif method >= 3:
            c64, r64, _ = ctx.bicgstab(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 100000, eps, ncases=1)
        else:
            c64, r64, _ = ctx.rbgs(fx, fa, fb, ffvhc.BC_P, 1.2, 100000, eps, ncases=1)
        ref = ctx.download(fx, cid, out=x.copy())
        scale = np.max(np.abs(ref[:, LIVE]))
        assert np.max(np.abs(got[:, LIVE] - ref[:, LIVE])) <= 1e-7 * scale
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


@pytest.mark.parametrize("name", ["two_level", "three_level"])
def test_coarse_operator_is_ptap(ctx, name):
    """The two-level preconditioner's coarse operator equals P^T A P by its definition (A applied
    by the oracle to each block's indicator, ghosts by the oracle's Updater1 with the preconditioned
    vectors' Neumann-0 outer BC, block sums), and the device's blocked Gauss-Jordan inverts it."""
    tree = trees()[name]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(31)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        fa, fb, fx = upload_system(ctx, cid, A, b, og.zeros())
        ac, ainv = ctx.coarse_operator(fa, cid)
        n = og.nb
        inner = (slice(G, G + N),) * 3
        sz = (ctypes.c_int * 3)(N, N, N)
        want = np.zeros((n, n))
        ax = np.zeros((N + 2 * G,) * 3)
        Ab = [[np.ascontiguousarray(A[q][a]) for q in range(7)] for a in range(n)]
        for blk in range(n):
            x = og.zeros()
            x[(blk,) + inner] = 1.0
            og.ghost_fill(x, 1, [1] * 6, [0.0] * 6)
            for a in range(n):
                og.L.orc_calc_ax(O.ptr(ax), *[O.ptr(Ab[a][q]) for q in range(7)],
                                 O.ptr(np.ascontiguousarray(x[a])), sz, G)
                want[a, blk] = ax[inner].sum()
        scale = np.max(np.abs(want))
        assert np.max(np.abs(ac - want)) <= 1e-12 * scale
        assert np.count_nonzero(np.abs(want) > 1e-12 * scale) > 2 * n  # couplings between blocks exist
        assert np.max(np.abs(ainv @ ac - np.eye(n))) <= 1e-8
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


@pytest.mark.parametrize("method", [None, 8, 19, 25])
def test_graph_replay_equals_launches(ctx, method):
    """BiCGSTAB with its iteration replayed as a captured HIP graph gives bit-identical results to
    launching the kernels one by one, over repeated solves (each solve captures and replays its
    own graph, with eager work in between)."""
    tree = trees()["three_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(41)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        fa, fb, fx = upload_system(ctx, cid, A, b, og.zeros())
        got = {}
        for graphs in (False, True, False, True):
            ctx.set_graphs(graphs)
            ctx.upload(fx, cid, og.zeros())
            if method is None:
                cnt, res, rc = ctx.bicgstab(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 5000, 1e-10, ncases=1)
            else:
                cnt, res, _o, rc = ctx.solve_mixed(method, fx, fa, fb, ffvhc.BC_P, 1.4, 5000, 1e-10, 0.5, ncases=1)
            assert rc == 0 and res[0] <= 1e-10
            x = ctx.download(fx, cid, out=og.zeros())
            got.setdefault(graphs, []).append((int(cnt[0]), x))
        ref_cnt, ref_x = got[False][0]
        for graphs in (False, True):
            for c_, x in got[graphs]:
                assert c_ == ref_cnt and np.array_equal(x, ref_x)
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.set_graphs(False)
        ctx.remove_case(cid)


@pytest.mark.parametrize("method", [None, 19, 26])
def test_launch_batch_hint_changes_nothing(ctx, method):
    """The first solve after a selection change launches growing batches (no hint); a repeated solve
    starts with one batch of the previous count (hinted).  Counts and every bit of x are equal."""
    tree = trees()["two_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(43)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        fa, fb, fx = upload_system(ctx, cid, A, b, og.zeros())
        ctx.select(None)
        ctx.select([cid])  # a new active set starts without launch hints
        got = []
        for _ in range(3):
            ctx.upload(fx, cid, og.zeros())
            if method is None:
                cnt, res, rc = ctx.bicgstab(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 5000, 1e-10, ncases=1)
            else:
                cnt, res, _o, rc = ctx.solve_mixed(method, fx, fa, fb, ffvhc.BC_P, 1.4, 5000, 1e-10, 0.5, ncases=1)
            assert rc == 0 and res[0] <= 1e-10
            got.append((int(cnt[0]), ctx.download(fx, cid, out=og.zeros())))
        for c_, x in got[1:]:
            assert c_ == got[0][0] and np.array_equal(x, got[0][1])
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.select(None)
        ctx.remove_case(cid)


@pytest.mark.parametrize("method", [23, 26, 28, 29])
def test_fused_finalisation_bit_identical(ctx, method):
    """The multiplicative two-level methods form each iteration's end (convergence test, next rho,
    beta) inside the next iteration's first coarse application, except after a batch's last
    iteration.  With that on (default) and off, two cases solved in one launch (converging at
    different iterations, so one finishes inside a batch while the other goes on) give the same
    counts, residuals and every bit of x; so does a solve whose inner cap ends every round."""
    trs = (trees()["two_level"], trees()["flat"])
    params = ffvhc.case_params(bc=BC)
    bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
    bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
    ids = [ctx.add_case(tr, params) for tr in trs]
    fa = [ctx.field(8) for _ in range(7)]
    fb, fx = ctx.field(8), ctx.field(8)
    try:
        ogs = [OracleGrid(tr) for tr in trs]
        for k in range(2):
            A, b = make_system(ogs[k], np.random.default_rng(60 + k), bct, bcv)
            for q in range(7):
                ctx.upload(fa[q], ids[k], A[q])
            ctx.upload(fb, ids[k], b)
        for cap, eps in ((5000, 1e-10), (6, 1e-10)):
            got = []
            for fuse in (True, False, True):
                ctx.set_fuse_fin(fuse)
                for k in range(2):
                    ctx.upload(fx, ids[k], ogs[k].zeros())
                inner = 1e-4 if method == 29 else 0.5  # 29: the FP32 inner solve's tolerance per correction
                cnt, res, outer, rc = ctx.solve_mixed(method, fx, fa, fb, ffvhc.BC_P, 1.4, cap, eps, inner, ncases=2)
                xs = [ctx.download(fx, ids[k], out=ogs[k].zeros()) for k in range(2)]
                got.append((list(cnt), list(res), list(outer), rc, xs))
            if cap == 5000:
                assert got[0][3] == 0 and max(got[0][1]) <= 1e-10
                assert got[0][0][0] != got[0][0][1]  # the cases stop at different iterations
            for g in got[1:]:
                assert g[0] == got[0][0] and g[1] == got[0][1] and g[2] == got[0][2] and g[3] == got[0][3]
                for k in range(2):
                    assert np.array_equal(g[4][k], got[0][4][k])
    finally:
        ctx.set_fuse_fin(True)
        for f in fa + [fb, fx]:
            ctx.free(f)
        for i in ids:
            ctx.remove_case(i)


def class_uniform_system(og, rng, bct, bcv, n_other=3):
    """Bands constant per block (a random factor per block, as 1/dx^2 is per level) with the outer
    BC folded in: every block is uniform per stencil position class.  Then one cell's band changes
    in n_other blocks, which makes those blocks non-uniform."""
    nb = og.shape[0]
    lb = rng.uniform(0.5, 2.0, nb)[:, None, None, None]
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = lb
    A[0][:] = -6.0 * lb - 0.05
    b = rng.uniform(-1, 1, og.shape)
    og.L.orc_matrix_bc(og.ref(), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv))
    other = rng.choice(nb, n_other, replace=False)
    for blk in other:
        i, j, k = (G + rng.integers(0, N, 3)).tolist()
        A[2][blk, i, j, k] *= 1.1
    return A, b, other


@pytest.mark.parametrize("method,n_other", [(15, 3), (19, 3), (26, 3), (19, 0), (26, 0)])
def test_stencil_classes_bit_identical(ctx, method, n_other):
    """The compact-record methods read the bands and records of class-uniform blocks from the 27
    class rows; with that on (default) and off the solve is bit-identical, on a system where every
    block is class-uniform (the class-only kernels) and on one where a few are not (the mixed
    kernels, which pick per block)."""
    tree = trees()["three_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(77)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b, other = class_uniform_system(og, rng, bct, bcv, n_other)
        fa, fb, fx = upload_system(ctx, cid, A, b, og.zeros())
        nu, no = ctx.stencil_classes(fa)
        assert (nu, no) == (og.shape[0] - len(other), len(other))
        got = []
        for classes in (True, False, True):
            ctx.set_classes(classes)
            ctx.upload(fx, cid, og.zeros())
            cnt, res, _o, rc = ctx.solve_mixed(method, fx, fa, fb, ffvhc.BC_P, 1.4, 5000, 1e-10, 0.5, ncases=1)
            assert rc == 0 and res[0] <= 1e-10
            got.append((int(cnt[0]), ctx.download(fx, cid, out=og.zeros())))
        for c_, x in got[1:]:
            assert c_ == got[0][0] and np.array_equal(x, got[0][1])
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.set_classes(True)
        ctx.remove_case(cid)


@pytest.mark.parametrize("method", [13, 19, 26])
def test_shadow_residual_choices(ctx, method):
    """The compact-record methods with the reference's r0 = r (default) and with the seeded hash
    shadow residual both reach the FP64 residual and agree; each is deterministic."""
    tree = trees()["two_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(5)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b, _ = class_uniform_system(og, rng, bct, bcv, 2)
        fa, fb, fx = upload_system(ctx, cid, A, b, og.zeros())
        got = {}
        for hashed in (True, False, True, False):
            ctx.set_shadow_hash(hashed)
            ctx.upload(fx, cid, og.zeros())
            cnt, res, _o, rc = ctx.solve_mixed(method, fx, fa, fb, ffvhc.BC_P, 1.4, 5000, 1e-10, 0.5, ncases=1)
            assert rc == 0 and res[0] <= 1e-10
            x = ctx.download(fx, cid, out=og.zeros())
            if hashed in got:
                assert got[hashed][0] == int(cnt[0]) and np.array_equal(got[hashed][1], x)
            got[hashed] = (int(cnt[0]), x)
        xa, xb = got[True][1][:, LIVE], got[False][1][:, LIVE]
        assert np.max(np.abs(xa - xb)) <= 1e-6 * np.max(np.abs(xa))
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.set_shadow_hash(False)
        ctx.remove_case(cid)


def run_oracle_bicgstab(og, x, A, b, bct, bcv, omega, cmax, eps):
    cnt, res = ctypes.c_int(), ctypes.c_double()
    sc = [og.zeros() for _ in range(9)]
    og.L.orc_bicgstab(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.parr(sc), O.cint6(bct), O.cdbl6(bcv),
                      ctypes.c_double(omega), 1, cmax, ctypes.c_double(eps), ctypes.byref(cnt), ctypes.byref(res))
    return cnt.value, res.value


@pytest.mark.parametrize("name", ["flat", "three_level"])
def test_bicgstab_fixed_count_matches_oracle(ctx, name):
    tree = trees()[name]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(12)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        x = og.zeros()
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.bicgstab(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 8, 0.0, ncases=1)
        got = ctx.download(fx, cid, out=x.copy())
        c2, r2 = run_oracle_bicgstab(og, x, A, b, bct, bcv, 1.0, 8, 0.0)
        assert rc == 0 and cnt[0] == c2 == 9
        assert abs(res[0] - r2) <= 1e-9 * abs(r2)
        scale = np.max(np.abs(x[:, LIVE]))
        assert np.max(np.abs(got[:, LIVE] - x[:, LIVE])) <= 1e-11 * scale
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)


def test_bicgstab_converges_and_mixed_matches(ctx):
    tree = trees()["three_level"]
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params)
    try:
        og = OracleGrid(tree)
        rng = np.random.default_rng(13)
        bct = [params.bct[ffvhc.BC_P][q] for q in range(6)]
        bcv = [params.bcv[ffvhc.BC_P][q] for q in range(6)]
        A, b = make_system(og, rng, bct, bcv)
        x = og.zeros()
        fa, fb, fx = upload_system(ctx, cid, A, b, x)
        cnt, res, rc = ctx.bicgstab(fx, fa, fb, ffvhc.BC_P, 1.0, 1, 5000, 1e-10, ncases=1)
        assert rc == 0 and res[0] <= 1e-10 and 0 < cnt[0] < 5000
        ref = ctx.download(fx, cid, out=x.copy())
        assert true_residual(og, A, ref, b) <= 1.001e-10
        ctx.upload(fx, cid, og.zeros())
        cm, rm, outer, rc = ctx.solve_mixed(2, fx, fa, fb, ffvhc.BC_P, 1.0, 5000, 1e-10, 1e-5, ncases=1)
        assert rc == 0 and rm[0] <= 1e-10 and outer[0] >= 2
        got = ctx.download(fx, cid, out=x.copy())
        assert true_residual(og, A, got, b) <= 1.001e-10
        scale = np.max(np.abs(ref[:, LIVE]))
        assert np.max(np.abs(got[:, LIVE] - ref[:, LIVE])) <= 1e-7 * scale
        for f in fa + [fb, fx]:
            ctx.free(f)
    finally:
        ctx.remove_case(cid)
