"""GPU parity of the fractional time step (Solver::Update with explicit U) against the oracle.

The device stage kernels evaluate the reference's per-cell expressions in the same order in
FP64 with no contraction, and RBGS is bit-exact for a fixed sweep count, so every field after
two steps must equal the oracle's bit for bit (interior cells and the first ghost layer).  The
Krylov variants differ only through reduction order: checked to 1e-10 relative.
Geometry is synthetic: seeded cut faces (float32 distances, ids 1-2) and solid cells.
"""
import ctypes

import numpy as np
import pytest

import ffvhc
from oracle import oracle as O
from helpers import G, N, OracleGrid, OracleStep, geometry, initial_arrays, live_mask

pytestmark = pytest.mark.gpu

LIVE = live_mask()
INTERIOR = (slice(G, G + N),) * 3

BC = {ffvhc.BC_UX: [(0, 1.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)],
      ffvhc.BC_UY: [(0, 0.0), (1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)],
      ffvhc.BC_UZ: [(0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (0, 0.0), (1, 0.0)],
      ffvhc.BC_P: [(1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)]}


def trees():
    return {
        "flat": ffvhc.Tree(root=(3, 2, 2), root_length=1.0),
        "two_level": ffvhc.Tree(root=(2, 2, 2), root_length=1.0,
                                regions=[dict(kind="ball", a=(1.0, 1.1, 0.9), r=0.35, level=1)]),
    }


def upload_all(ctx, cid, F, a):
    ctx.upload(F.ux, cid, a["ux"])
    ctx.upload(F.uy, cid, a["uy"])
    ctx.upload(F.uz, cid, a["uz"])
    ctx.upload(F.p, cid, a["p"])
    for q in range(6):
        ctx.upload(F.v[q], cid, a["v"][q])


def compare(ctx, cid, F, a, tol, what):
    pairs = [("ux", F.ux, a["ux"]), ("uy", F.uy, a["uy"]), ("uz", F.uz, a["uz"]), ("p", F.p, a["p"]),
             ("lapp", F.lapp, a["lapp"]), ("b", F.b, a["b"])]
    pairs += [("v%d" % q, F.v[q], a["v"][q]) for q in range(6)]
    pairs += [("A%d" % q, F.A[q], a["A"][q]) for q in range(7)]
    # after the step the oracle's ucp/udp equal uc/ud; the device's ucp/udp handles hold them
    pairs += [("uc%d" % d, F.ucp[d], a["uc"][d]) for d in range(3)]
    pairs += [("ud%d" % d, F.udp[d], a["ud"][d]) for d in range(3)]
    for name, fld, want in pairs:
        got = ctx.download(fld, cid, out=np.zeros_like(want))
        mask = LIVE if name in ("ux", "uy", "uz", "p", "lapp") else np.zeros_like(LIVE)
        if name not in ("ux", "uy", "uz", "p", "lapp"):
            mask[INTERIOR] = True
        g, w = got[:, mask], want[:, mask]
        if tol == 0.0:
            bad = np.flatnonzero(g != w)
            assert bad.size == 0, "%s %s: %d cells differ, max |d| %.3e" % (what, name, bad.size,
                                                                           np.max(np.abs(g - w)))
        else:
            scale = max(np.max(np.abs(w)), 1e-300)
            err = np.max(np.abs(g - w)) / scale
            assert err <= tol, "%s %s: rel err %.3e" % (what, name, err)


@pytest.fixture(scope="module")
def ctx():
    c = ffvhc.Context(max_blocks=1024)
    yield c
    c.close()


def run_case(ctx, name, solver, omega, count_max, eps, tol, seed=5, nsteps=2, implicit=None, tree=None,
             bc=None, params_kw=None, ref=None, forcing=False, geo=None):
    """n steps of the device and of the oracle from the same seeded state; every field compared.
    tree / bc / params_kw override the synthetic case; ref = (x, y, z, p) turns on
    bcut_set_reference_value; forcing uploads a seeded temperature deviation and body force (their
    terms act through gravity/betag and f); geo = (c, cid, pid) replaces the seeded cut cells."""
    tree = tree if tree is not None else trees()[name]
    og = OracleGrid(tree)
    rng = np.random.default_rng(seed)
    geo = geo if geo is not None else geometry(og, rng)
    params = ffvhc.case_params(**dict(dict(rho=1.0, mu=0.01, dt=0.002, bc=bc or BC), **(params_kw or {})))
    cid = ctx.add_case(tree, params, geometry=(geo[0], geo[1], geo[2]))
    F = ctx.step_fields(with_t=forcing, with_f=forcing)
    try:
        a = initial_arrays(og, rng)
        if forcing:
            a["t"][:] = rng.standard_normal(og.shape)
            for d in range(3):
                a["f"][d][:] = rng.standard_normal(og.shape)
            ctx.upload(F.t, cid, a["t"])
            for d in range(3):
                ctx.upload(F.f[d], cid, a["f"][d])
        # the oracle starts from ghost-filled fields, as Solver's initial fills leave them
        bl = lambda grp: ([params.bct[grp][f] for f in range(6)], [params.bcv[grp][f] for f in range(6)])
        for d, key in enumerate(("ux", "uy", "uz")):
            og.ghost_fill(a[key], 5, *bl(d))
        og.ghost_fill(a["p"], 1, *bl(ffvhc.BC_P))
        upload_all(ctx, cid, F, a)
        orc = OracleStep(og, params, geo, a, solver, omega, count_max, eps, implicit)
        if ref is not None:
            orc.P.ref_active = 1
            orc.P.ref_x, orc.P.ref_y, orc.P.ref_z, orc.P.ref_p = ref
        ctx.select([cid])
        kw = {}
        if implicit is not None:
            kw = dict(implicit_u=True, alpha_u=implicit[0], omega_u=implicit[1], count_max_u=implicit[3],
                      eps_u=implicit[2])
        if ref is not None:
            kw["ref"] = ref
        for n in range(nsteps):
            info_o = orc.step(n)
            info_d, ms, rc = ctx.step(F, solver=solver, omega=omega, count_max=count_max, eps=eps, ncases=1, **kw)
            assert rc == 0 and info_o.status == 0
            assert info_d[0].countP == info_o.countP, (info_d[0].countP, info_o.countP)
            if implicit is not None:
                assert list(info_d[0].countU) == list(info_o.countU)
            compare(ctx, cid, F, a, tol, "%s step %d" % (name, n))
    finally:
        ctx.select(None)
        extra = ([F.t] + list(F.f)) if forcing else []
        for fld in ([F.ux, F.uy, F.uz, F.p, F.b, F.lapp] + list(F.uc) + list(F.ucp) + list(F.ud) + list(F.udp) +
                    list(F.v) + list(F.A) + extra):
            ctx.free(fld)
        ctx.remove_case(cid)


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_step_rbgs_bit_exact(ctx, name):
    run_case(ctx, name, solver=0, omega=1.2, count_max=25, eps=1e-30, tol=0.0)


@pytest.mark.parametrize("name", ["two_level"])
def test_step_pcg_matches(ctx, name):
    run_case(ctx, name, solver=2, omega=1.0, count_max=15, eps=1e-30, tol=1e-10)


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_step_implicit_u_matches(ctx, name):
    """UpdateUX/UY/UZ: calc_abd_u (bit-exact per cell) + BiCGSTAB per component (reduction order
    differs) then the RBGS pressure solve."""
    run_case(ctx, name, solver=0, omega=1.2, count_max=20, eps=1e-30, tol=1e-9, implicit=(0.5, 1.0, 1e-30, 8))


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_step_reference_pressure_bit_exact(ctx, name):
    """bcut_set_reference_value (bcut.f90:5626-5710, with its Ae(i+1) index) pins p at a point:
    the rows of the cell holding it and of its six neighbours are rewritten after calc_ab_p."""
    ref = (0.53, 0.41, 0.37, 0.25) if name == "flat" else (0.97, 1.06, 0.84, -0.1)
    run_case(ctx, name, solver=0, omega=1.2, count_max=25, eps=1e-30, tol=0.0, ref=ref)


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_step_gravity_buoyancy_body_force_bit_exact(ctx, name):
    """Nonzero gravity in the predictor and bcut_add_g (with the buoyancy term betag (T - Tr) on
    z), and a body force through bcut_add_f (bcut.f90:3793-3877)."""
    run_case(ctx, name, solver=0, omega=1.2, count_max=25, eps=1e-30, tol=0.0, forcing=True,
             params_kw=dict(gravity=(0.3, -0.7, -9.8), betag=0.02, tr=0.4))


def test_step_periodic_outer_bc_bit_exact(ctx):
    """Periodic outer faces (type 9, bcx.f90:585-658): a tree one block thick in y and z, every
    field periodic there; the plus-side call writes both ghost layers, the minus side is a no-op."""
    tree = ffvhc.Tree(root=(3, 1, 1), root_length=1.0)
    per = (9, 0.0)
    bc = {g: faces[:2] + [per] * 4 for g, faces in BC.items()}
    run_case(ctx, "periodic", solver=0, omega=1.2, count_max=25, eps=1e-30, tol=0.0, tree=tree, bc=bc)


def test_step_three_level_bit_exact(ctx):
    """A time step on a three-level tree (level jumps of one at every interface)."""
    tree = ffvhc.Tree(root=(2, 2, 2), root_length=1.0,
                      regions=[dict(kind="ball", a=(0.9, 1.1, 1.05), r=0.2, level=2)])
    assert tree.max_level == 2
    run_case(ctx, "three_level", solver=0, omega=1.2, count_max=25, eps=1e-30, tol=0.0, tree=tree)


def test_step_bicgstab_pressure_matches(ctx):
    """SolvingMethod/P = BiCGSTAB (the reference example's choice): fixed count, reduction order
    differs from the oracle."""
    run_case(ctx, "two_level", solver=3, omega=1.0, count_max=12, eps=1e-30, tol=1e-10)


def c1_case(radius=0.5):
    """The flow past a sphere of the single-level configuration: 8x4x4 roots of 16^3 cells, root
    length 2 (h = 1/8), sphere of diameter 1 at (4, 4, 4), cut cells from the analytic sphere."""
    tree = ffvhc.Tree(root=(8, 4, 4), root_length=2.0)
    c, cid, pid, counts = tree.geometry_sphere(center=(4.0, 4.0, 4.0), radius=radius, cut_id=1)
    return tree, (c, cid, pid), counts


def test_step_c1_sphere_geometry_bit_exact(ctx):
    """Two steps of the reference's single-level flow past a sphere (Re 100, dt 0.01, RBGS omega
    1.2, fixed sweep count) with the sphere's cut cells: every field equals the oracle's."""
    tree, geo, counts = c1_case()
    assert counts[1] > 0
    bc = {ffvhc.BC_UX: [(0, 1.0)] + [(1, 0.0)] * 5, ffvhc.BC_UY: [(0, 0.0)] + [(1, 0.0)] * 5,
          ffvhc.BC_UZ: [(0, 0.0)] + [(1, 0.0)] * 5, ffvhc.BC_P: [(1, 0.0), (0, 0.0)] + [(1, 0.0)] * 4}
    run_case(ctx, "c1", solver=0, omega=1.2, count_max=12, eps=1e-30, tol=0.0, tree=tree, bc=bc, geo=geo,
             params_kw=dict(mu=0.01, dt=0.01))


def test_step_three_level_sphere_wake_bit_exact(ctx):
    """The three-level sphere + wake configuration (SphereDivider3-style refinement, Re 300, dt at
    CFL 0.3 on the finest cells) at reduced size: roots 6x3x3, sphere of diameter 0.3, 516 blocks
    on levels 0-2 with the sphere's cut cells; two steps equal the oracle's bit for bit."""
    import flows
    tree, geo, counts = flows.three_level_sphere(root=(6, 3, 3), root_length=1.0, d=0.3)
    assert tree.max_level == 2 and counts[1] > 0
    lv = np.bincount(tree.table()["level"], minlength=3)
    assert np.all(lv > 0), lv
    run_case(ctx, "c3_reduced", solver=0, omega=1.2, count_max=10, eps=1e-30, tol=0.0, tree=tree,
             bc=flows.FLOW_BC, geo=geo, params_kw=dict(mu=1.0 * 0.3 / 300.0, dt=flows.cfl_dt(tree, 0.3)))


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_force_p_matches_oracle(ctx, name):
    """bcut_calc_f_p summed over the blocks (Solver::PrintForceCID) for each boundary id."""
    tree = trees()[name]
    og = OracleGrid(tree)
    rng = np.random.default_rng(21)
    geo = geometry(og, rng)
    params = ffvhc.case_params(bc=BC)
    cid = ctx.add_case(tree, params, geometry=(geo[0], geo[1], geo[2]))
    pf = ctx.field(8)
    try:
        p = rng.standard_normal(og.shape)
        og.ghost_fill(p, 1, [params.bct[ffvhc.BC_P][f] for f in range(6)],
                      [params.bcv[ffvhc.BC_P][f] for f in range(6)])
        ctx.upload(pf, cid, p)
        ctx.select([cid])
        sz = (ctypes.c_int * 3)(N, N, N)
        for target in (1, 2):
            got = ctx.force_p(pf, target, ncases=1)[0]
            want = np.zeros(3)
            npad = (N + 2 * G) ** 3
            for b in range(og.nb):
                fs = [np.zeros(npad) for _ in range(3)]
                out = np.zeros(3)
                cb = [np.ascontiguousarray(geo[0][f][b]) for f in range(6)]
                ib = [np.ascontiguousarray(geo[1][f][b]) for f in range(6)]
                og.L.orc_bcut_calc_f_p(O.ptr(fs[0]), O.ptr(fs[1]), O.ptr(fs[2]), O.ptr(out), target,
                                       O.ptr(np.ascontiguousarray(p[b])), O.parr(cb), O.parr(ib, ctypes.c_int),
                                       O.ptr(np.ascontiguousarray(geo[2][b])), ctypes.c_double(og.tb["dx"][b]),
                                       ctypes.c_double(0.01), sz, G)
                want += out
            assert np.all(np.abs(got - want) <= 1e-12 * np.max(np.abs(want)) + 1e-300), (got, want)
    finally:
        ctx.select(None)
        ctx.free(pf)
        ctx.remove_case(cid)


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_force_v_matches_oracle(ctx, name):
    """bcut_calc_f_v summed over the blocks (Solver::PrintForceCID) for each boundary id, with the
    case's mu and velocity ghosts filled by Scalar3DUpdater5."""
    tree = trees()[name]
    og = OracleGrid(tree)
    rng = np.random.default_rng(22)
    geo = geometry(og, rng)
    params = ffvhc.case_params(mu=0.013, bc=BC)
    cid = ctx.add_case(tree, params, geometry=(geo[0], geo[1], geo[2]))
    uf = [ctx.field(8) for _ in range(3)]
    try:
        u = [rng.standard_normal(og.shape) for _ in range(3)]
        for d in range(3):
            og.ghost_fill(u[d], 5, [params.bct[d][f] for f in range(6)], [params.bcv[d][f] for f in range(6)])
            ctx.upload(uf[d], cid, u[d])
        ctx.select([cid])
        sz = (ctypes.c_int * 3)(N, N, N)
        for target in (1, 2):
            got = ctx.force_v(uf[0], uf[1], uf[2], target, ncases=1)[0]
            want = np.zeros(3)
            npad = (N + 2 * G) ** 3
            for b in range(og.nb):
                fs = [np.zeros(npad) for _ in range(3)]
                out = np.zeros(3)
                cb = [np.ascontiguousarray(geo[0][f][b]) for f in range(6)]
                ib = [np.ascontiguousarray(geo[1][f][b]) for f in range(6)]
                ub = [np.ascontiguousarray(u[d][b]) for d in range(3)]
                og.L.orc_bcut_calc_f_v(O.ptr(fs[0]), O.ptr(fs[1]), O.ptr(fs[2]), O.ptr(out), target,
                                       O.ptr(ub[0]), O.ptr(ub[1]), O.ptr(ub[2]), O.parr(cb),
                                       O.parr(ib, ctypes.c_int), O.ptr(np.ascontiguousarray(geo[2][b])),
                                       ctypes.c_double(params.rho), ctypes.c_double(params.mu),
                                       ctypes.c_double(og.tb["dx"][b]), ctypes.c_double(0.01),
                                       ctypes.c_double(0.0), sz, G)
                want += out
            assert np.any(want != 0.0)
            # per-block order differs from the reference's serial loop: rounding-level tolerance
            assert np.all(np.abs(got - want) <= 1e-12 * np.max(np.abs(want)) + 1e-300), (got, want)
    finally:
        ctx.select(None)
        for f in uf:
            ctx.free(f)
        ctx.remove_case(cid)


@pytest.mark.parametrize("method", [15, 19, 26])
def test_step_with_fast_poisson_matches_fp64(ctx, method):
    """Two time steps of a flow past a cut-cell sphere: the pressure solved by the FP32-matrix
    BiCGSTAB with block SOR (+ coarse correction) gives the same fields as the FP64 BiCGSTAB at
    the same 1e-10 residual, within the solver tolerance; the cut-cell matrix is not FP32-exact, so
    the FP64 acceptance check takes part."""
    from serve import uniform_flow_init
    tree = ffvhc.Tree(root=(4, 2, 2), root_length=2.0)
    c, cid, pid, counts = tree.geometry_sphere(center=(2.0, 2.0, 2.0), radius=0.5, cut_id=1)
    assert counts[1] > 0
    bc = {ffvhc.BC_UX: [(0, 1.0)] + [(1, 0.0)] * 5, ffvhc.BC_UY: [(0, 0.0)] + [(1, 0.0)] * 5,
          ffvhc.BC_UZ: [(0, 0.0)] + [(1, 0.0)] * 5, ffvhc.BC_P: [(1, 0.0), (0, 0.0)] + [(1, 0.0)] * 4}
    params = ffvhc.case_params(mu=0.01, dt=0.01, bc=bc)
    out = []
    for kw in (dict(solver=3, omega=1.0, count_max=5000, eps=1e-10),
               dict(solver=3, omega=1.6, count_max=5000, eps=1e-10, mixed=True, mixed_method=method)):
        case = ctx.add_case(tree, params, geometry=(c, cid, pid))
        F = ctx.step_fields()
        try:
            ctx.select([case])
            uniform_flow_init(0, u0=1.0, noise=0.0)(ctx, case, F)
            for _ in range(2):
                info, _ms, rc = ctx.step(F, ncases=1, **kw)
                assert rc == 0 and info[0].residualP <= 1e-10
            shp = (tree.nblk, 20, 20, 20)
            out.append([ctx.download(f, case, out=np.zeros(shp)) for f in (F.ux, F.uy, F.uz, F.p)])
        finally:
            ctx.select(None)
            for f in [F.ux, F.uy, F.uz, F.p, F.b, F.lapp] + list(F.uc) + list(F.ucp) + list(F.ud) + list(F.udp) + \
                    list(F.v) + list(F.A):
                ctx.free(f)
            ctx.remove_case(case)
    inner = (slice(None),) + (slice(2, 18),) * 3
    for a, b in zip(out[0], out[1]):
        scale = np.max(np.abs(a[inner]))
        assert np.max(np.abs(a[inner] - b[inner])) <= 1e-6 * scale


def test_step_setup_reuse_bit_identical(ctx):
    """The time step keeps the setup of its mixed-precision pressure solve from step to step
    (FFVHC_OPTION_SETUP_REUSE): three steps of a flow past a sphere, then the case replaced by
    one with a smaller sphere in the same block slots, three more steps (the kept setup must not
    be used for the other geometry), and a standalone solve between two steps (which writes the
    setup buffers): every field is the same bit for bit as with the setup made every step."""
    from serve import uniform_flow_init
    tree = ffvhc.Tree(root=(4, 2, 2), root_length=2.0)
    bc = {ffvhc.BC_UX: [(0, 1.0)] + [(1, 0.0)] * 5, ffvhc.BC_UY: [(0, 0.0)] + [(1, 0.0)] * 5,
          ffvhc.BC_UZ: [(0, 0.0)] + [(1, 0.0)] * 5, ffvhc.BC_P: [(1, 0.0), (0, 0.0)] + [(1, 0.0)] * 4}
    params = ffvhc.case_params(mu=0.01, dt=0.01, bc=bc)
    kw = dict(solver=3, omega=1.6, count_max=5000, eps=1e-10, mixed=True, mixed_method=26)
    shp = (tree.nblk, 20, 20, 20)
    runs = []
    for reuse in (False, True):
        ctx.set_setup_reuse(reuse)
        F = ctx.step_fields()
        out = []
        try:
            for radius in (0.5, 0.4):
                c, cid, pid, counts = tree.geometry_sphere(center=(2.0, 2.0, 2.0), radius=radius, cut_id=1)
                case = ctx.add_case(tree, params, geometry=(c, cid, pid))
                try:
                    ctx.select([case])
                    uniform_flow_init(0, u0=1.0, noise=0.0)(ctx, case, F)
                    for k in range(3):
                        info, _ms, rc = ctx.step(F, ncases=1, **kw)
                        assert rc == 0 and info[0].residualP <= 1e-10
                        if k == 1 and radius == 0.4:
                            # a solve of another matrix (2 A) writes the setup buffers in between
                            A2 = [ctx.field(8) for _ in range(7)]
                            x = ctx.field(8)
                            try:
                                for f, g in zip(F.A, A2):
                                    ctx.upload(g, case, 2.0 * ctx.download(f, case, out=np.zeros(shp)))
                                ctx.fill(x, 0.0)
                                *_, st = ctx.solve_mixed(26, x, A2, F.b, ffvhc.BC_P, 1.6, 50, 1e-6, 0.5)
                                assert st == 0
                            finally:
                                for f in A2 + [x]:
                                    ctx.free(f)
                    out += [ctx.download(f, case, out=np.zeros(shp)) for f in (F.ux, F.uy, F.uz, F.p)]
                finally:
                    ctx.select(None)
                    ctx.remove_case(case)
        finally:
            ctx.set_setup_reuse(True)
            for f in [F.ux, F.uy, F.uz, F.p, F.b, F.lapp] + list(F.uc) + list(F.ucp) + list(F.ud) + list(F.udp) + \
                    list(F.v) + list(F.A):
                ctx.free(f)
        runs.append(out)
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("method", [19, 26])
def test_step_setup_reuse_follows_reference_point_and_selection(ctx, method):
    """The kept pressure setup must be made again when the reference-pressure point is switched
    on, moved or switched off (bcut_set_reference_value rewrites matrix rows), and stays valid
    across a selection that changes and comes back: two resident flow cases, a schedule of steps
    with ref off / at one point / at another / off, the second case selected in and out; every
    field bit-identical with the setup kept (FFVHC_OPTION_SETUP_REUSE 1) and made every step (0),
    for an additive (19) and a multiplicative (26) two-level method."""
    from serve import uniform_flow_init
    tree = ffvhc.Tree(root=(4, 2, 2), root_length=2.0)
    bc = {ffvhc.BC_UX: [(0, 1.0)] + [(1, 0.0)] * 5, ffvhc.BC_UY: [(0, 0.0)] + [(1, 0.0)] * 5,
          ffvhc.BC_UZ: [(0, 0.0)] + [(1, 0.0)] * 5, ffvhc.BC_P: [(1, 0.0), (0, 0.0)] + [(1, 0.0)] * 4}
    params = ffvhc.case_params(mu=0.01, dt=0.01, bc=bc)
    kw = dict(solver=3, omega=1.6, count_max=5000, eps=1e-10, mixed=True, mixed_method=method)
    schedule = [(None, 1), (None, 2), ((1.1, 2.3, 1.7, 0.0), 1), ((1.1, 2.3, 1.7, 0.0), 2),
                ((3.2, 1.4, 2.6, 0.1), 2), ((3.2, 1.4, 2.6, 0.1), 1), (None, 1)]
    shp = (tree.nblk, 20, 20, 20)
    runs = []
    for reuse in (False, True):
        ctx.set_setup_reuse(reuse)
        cases, Fs = [], []
        try:
            for radius in (0.5, 0.4):
                c, cid, pid, _ = tree.geometry_sphere(center=(2.0, 2.0, 2.0), radius=radius, cut_id=1)
                cases.append(ctx.add_case(tree, params, geometry=(c, cid, pid)))
                Fs.append(ctx.step_fields())
                ctx.select([cases[-1]])
                uniform_flow_init(0, u0=1.0, noise=0.0)(ctx, cases[-1], Fs[-1])
            out = []
            for ref, nsel in schedule:
                for k in range(nsel):
                    ctx.select([cases[k]])
                    info, _ms, rc = ctx.step(Fs[k], ncases=1, ref=ref, **kw)
                    assert rc == 0 and info[0].residualP <= 1e-10
                for k in range(2):
                    out += [ctx.download(f, cases[k], out=np.zeros(shp)) for f in (Fs[k].ux, Fs[k].p)]
            runs.append(out)
        finally:
            ctx.set_setup_reuse(True)
            ctx.select(None)
            for F in Fs:
                for f in [F.ux, F.uy, F.uz, F.p, F.b, F.lapp] + list(F.uc) + list(F.ucp) + list(F.ud) + \
                        list(F.udp) + list(F.v) + list(F.A):
                    ctx.free(f)
            for cid in cases:
                ctx.remove_case(cid)
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("updater,group", [(1, ffvhc.BC_P), (5, ffvhc.BC_UX)])
def test_dump2_ghost_layers_match_oracle(ctx, tmp_path, updater, group):
    """Dump2 writes what ImposeBoundaryCondition leaves in both ghost layers: the file's padded
    blocks equal the oracle's ghost fill (vc = 2; same level, F2C, C2F, outer D/N BCs over the whole
    padded face) applied to the same interiors, on a three-level tree, bit for bit.  Cells the
    update never writes (block edges and corners away from outer faces) are 0 in both."""
    tree = trees()["two_level"] if updater == 1 else ffvhc.Tree(
        root=(2, 2, 2), root_length=1.0, regions=[dict(kind="ball", a=(0.9, 1.1, 1.05), r=0.2, level=2)])
    og = OracleGrid(tree)
    # Dirichlet and Neumann faces with nonzero values
    params = ffvhc.case_params(bc={group: [(0, 1.0), (1, 0.25), (0, -0.5), (1, 0.0), (1, 0.1), (0, 0.0)]})
    cid = ctx.add_case(tree, params)
    f1 = ctx.field(8)
    try:
        x = np.zeros(og.shape)
        x[(slice(None),) + INTERIOR] = np.random.default_rng(13).standard_normal((og.nb, N, N, N))
        ctx.upload(f1, cid, x)
        path = tmp_path / "pX-00000-0000000001.bin"
        ctx.dump(f1, cid, path, group=group, updater=updater)
        raw = np.fromfile(path, dtype=np.float64, offset=24).reshape(og.shape)
        want = x.copy()
        og.ghost_fill(want, updater, [params.bct[group][f] for f in range(6)],
                      [params.bcv[group][f] for f in range(6)])
        assert np.array_equal(raw, want)
        # the second layer is really written (not left at 0) on the outer faces and between blocks
        assert np.count_nonzero(raw[:, 0, G:G + N, G:G + N]) > 0 and np.count_nonzero(raw[:, G:G + N, G:G + N, 0]) > 0
    finally:
        ctx.free(f1)
        ctx.remove_case(cid)


def test_checkpoint_round_trip(ctx, tmp_path):
    """Dump2 of a field, then Load2 into another: the interior is identical, and after the ghost
    fill of both (Load2 ends with ImposeBoundaryCondition) so is the first ghost layer; the file has
    the reference's layout (header + padded blocks)."""
    tree = trees()["two_level"]
    og = OracleGrid(tree)
    cid = ctx.add_case(tree, ffvhc.case_params(bc=BC))
    f1, f2 = ctx.field(8), ctx.field(8)
    try:
        x = np.random.default_rng(3).standard_normal(og.shape)
        ctx.upload(f1, cid, x)
        path = tmp_path / "pP-00000-0000000007.bin"
        ctx.dump(f1, cid, path)
        head = ffvhc.dump_header(path)
        assert list(head) == [N, N, N, G, 1, og.nb]
        raw = np.fromfile(path, dtype=np.float64, offset=24).reshape(og.shape)
        inner = (slice(None),) + INTERIOR
        assert np.array_equal(raw[inner], x[inner])
        ctx.load(f2, cid, path)
        ctx.ghost_fill(f1, ffvhc.BC_P, 1)
        ctx.ghost_fill(f2, ffvhc.BC_P, 1)
        got = ctx.download(f2, cid, out=np.zeros(og.shape))
        ref = ctx.download(f1, cid, out=np.zeros(og.shape))
        assert np.array_equal(got[:, LIVE], ref[:, LIVE])
        assert np.array_equal(raw[:, LIVE], ref[:, LIVE])  # the file's first ghost layer is the fill's
    finally:
        ctx.free(f1)
        ctx.free(f2)
        ctx.remove_case(cid)


def test_load2_half_resolution(ctx, tmp_path):
    """Load2 of a file written for blocks of half the cell count (8^3, LocalScalar3D.cpp:313-331):
    every fine cell takes its parent coarse cell, x(i, j, k) = xc((i+1)/2, (j+1)/2, (k+1)/2) with
    1-based indices and integer division (sf3d_copy_x2, bsf3d.f90:5-33)."""
    tree = trees()["two_level"]
    og = OracleGrid(tree)
    cid = ctx.add_case(tree, ffvhc.case_params(bc=BC))
    f = ctx.field(8)
    try:
        nc, g = N // 2, G
        xc = np.random.default_rng(5).standard_normal((og.nb,) + (nc + 2 * g,) * 3)
        path = tmp_path / "pP-00000-0000000003.bin"
        with open(path, "wb") as fh:
            np.array([nc, nc, nc, g, 1, og.nb], np.int32).tofile(fh)
            xc.tofile(fh)
        ctx.load(f, cid, path)
        got = ctx.download(f, cid, out=np.zeros(og.shape))
        want = np.empty((og.nb, N, N, N))
        for k in range(1, N + 1):          # the reference's loop, 1-based, padded offset g - 1
            for j in range(1, N + 1):
                for i in range(1, N + 1):
                    want[:, k - 1, j - 1, i - 1] = xc[:, (k + 1) // 2 + g - 1, (j + 1) // 2 + g - 1,
                                                      (i + 1) // 2 + g - 1]
        assert np.array_equal(got[:, G:G + N, G:G + N, G:G + N], want)
        # a file whose blocks are neither 16^3 nor 8^3 is refused
        bad = tmp_path / "bad.bin"
        with open(bad, "wb") as fh:
            np.array([4, 4, 4, g, 1, og.nb], np.int32).tofile(fh)
            np.zeros((og.nb, 8, 8, 8)).tofile(fh)
        with pytest.raises(Exception):
            ctx.load(f, cid, bad)
    finally:
        ctx.free(f)
        ctx.remove_case(cid)


def test_field_stats_and_cfl_guard(ctx):
    """CalcStats per case against numpy on the same interior; the CFL guard of PrintStats."""
    tree = trees()["two_level"]
    og = OracleGrid(tree)
    cid = ctx.add_case(tree, ffvhc.case_params(dt=0.01, bc=BC))
    f = ctx.field(8)
    try:
        x = np.random.default_rng(4).standard_normal(og.shape)
        ctx.upload(f, cid, x)
        ctx.select([cid])
        got = ctx.stats(f, ncases=1)[0]
        xi = x[:, G:G + N, G:G + N, G:G + N]
        assert abs(got[0] - xi.sum()) <= 1e-12 * np.abs(xi).sum()
        assert got[1] == xi.max() and got[2] == xi.min()
        assert got[3] == np.abs(xi).max() and got[4] == np.abs(xi).min()
        dxl = 1.0 / 2 ** tree.max_level
        cfl, rc = ctx.cfl(f, ncases=1)
        assert abs(cfl[0] - np.abs(xi).max() * 0.01 / dxl) <= 1e-15 * cfl[0] and rc == 0
        ctx.upload(f, cid, x * 1e3)
        cfl, rc = ctx.cfl(f, ncases=1)
        assert rc == 2
    finally:
        ctx.select(None)
        ctx.free(f)
        ctx.remove_case(cid)
