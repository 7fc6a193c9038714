"""Grid-level oracle pins by mathematics: each grid-level oracle function is compared with a
textbook statement of its algorithm written here in numpy (or a library routine), or with an
invariant the discretisation fixes, and each pin is shown to fail when a mistake is planted in
the oracle (orc_plant, switched on only inside these tests).

* orc_bicgstab (FFVILS.h:793-974) vs textbook right-preconditioned BiCGSTAB (van der Vorst 1992,
  M^-1 = omega D^-1, the reference's one damped-Jacobi sweep from zero) and scipy's bicgstab
  (the same right-preconditioned form), iterate by iterate on flat, two- and three-level trees.
  The operator is A applied to the ghost-filled vector (numpy stencil; the ghost fill is pinned
  separately by linear fields in test_oracle_grid.py).
* orc_jacobi (FFVILS.h:49-102, 492-545) vs textbook damped Jacobi on the assembled operator.
* orc_rbgs_fixed_average (FFVILS.h:596-659) vs textbook red-black SOR + mean removal on one
  block; and through it a manufactured Neumann problem p = cos(pi x) cos(pi y) cos(pi z)
  converges at O(h^2) on one and eight blocks.
* orc_update (Solver.cpp:3344-3416, 5294-5771): a uniform flow with p = 0 is left unchanged
  exactly; a uniform flow with a random initial p is restored (remove_p adds back the predictor's
  gradient); after the projection the divergence of the face velocities equals the Poisson
  residual cell by cell (b - A p) and is at the solver tolerance.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import ffvhc
from oracle import oracle as O
from helpers import G, N, OracleGrid, OracleStep, geometry, interior

INT = (slice(None),) + (slice(G, G + N),) * 3
NULL_BC = ([1] * 6, [0.0] * 6)


@pytest.fixture
def plant():
    L = O.lib()

    def set_(i):
        L.orc_plant(i)

    yield set_
    L.orc_plant(0)


def _trees():
    return {
        "flat": ffvhc.Tree(root=(3, 2, 2), root_length=1.0),
        "two_level": ffvhc.Tree(root=(3, 3, 2), root_length=1.0,
                                regions=[dict(kind="ball", a=(1.4, 1.6, 1.0), r=0.45, level=1)]),
        "three_level": ffvhc.Tree(root=(2, 2, 2), root_length=1.0,
                                  regions=[dict(kind="ball", a=(0.9, 1.1, 1.05), r=0.2, level=2)]),
    }


TREES = _trees()


def stencil(A, x):
    """sum_q A_q x_nb(q) over the interior cells of padded arrays (bils.f90 calc_ax band order:
    Ap x, Aw x(i-1), Ae x(i+1), As x(j-1), An x(j+1), Ab x(k-1), At x(k+1); arrays are (b,k,j,i))."""
    def s(dk, dj, di):
        return x[:, G + dk:G + N + dk, G + dj:G + N + dj, G + di:G + N + di]
    Ai = [interior(a) for a in A]
    return (Ai[0] * s(0, 0, 0) + Ai[1] * s(0, 0, -1) + Ai[2] * s(0, 0, 1) + Ai[3] * s(0, -1, 0) +
            Ai[4] * s(0, 1, 0) + Ai[5] * s(-1, 0, 0) + Ai[6] * s(1, 0, 0))


class Operator:
    """y = A G(x) on the interior cells, G = the ghost fill of a scratch field (Neumann-0 outer
    ghosts, Scalar3DUpdater1 at level interfaces)."""

    def __init__(self, og, A, bc=NULL_BC):
        self.og, self.A, self.bc = og, A, bc
        self.n = og.nb * N ** 3

    def pad(self, v):
        x = self.og.zeros()
        x[INT] = v.reshape(self.og.nb, N, N, N)
        self.og.ghost_fill(x, 1, *self.bc)
        return x

    def __call__(self, v):
        return stencil(self.A, self.pad(v)).ravel()

    def linop(self):
        return spla.LinearOperator((self.n, self.n), matvec=self, dtype=np.float64)


def system(og, seed, bct):
    """Random diagonally dominant non-symmetric 7-band system with its outer BC folded in."""
    rng = np.random.default_rng(seed)
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = rng.uniform(0.5, 1.5, og.shape)
    A[0][:] = -(sum(A[q] for q in range(1, 7))) - rng.uniform(0.0, 0.5, og.shape)
    b = og.zeros()
    b[INT] = rng.uniform(-1, 1, (og.nb, N, N, N))
    og.L.orc_matrix_bc(og.ref(), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6([0.0] * 6))
    return A, b


def textbook_bicgstab(op, dinv, b, iters, omega):
    """Right-preconditioned BiCGSTAB (van der Vorst 1992), shadow residual r^ = r_0, x_0 = 0,
    M^-1 v = omega D^-1 v.  Returns x and the recursive residual after `iters` iterations."""
    x = np.zeros_like(b)
    r = b - op(x)
    rh = r.copy()
    rho_prev = alpha = w = 1.0
    p = v = np.zeros_like(b)
    for i in range(iters):
        rho = rh @ r
        if i == 0:
            p = r.copy()
        else:
            p = r + (rho / rho_prev) * (alpha / w) * (p - w * v)
        ph = omega * dinv * p
        v = op(ph)
        alpha = rho / (rh @ v)
        s = r - alpha * v
        sh = omega * dinv * s
        t = op(sh)
        w = (t @ s) / (t @ t)
        x = x + alpha * ph + w * sh
        r = s - w * t
        rho_prev = rho
    return x, r


def oracle_bicgstab(og, A, b, bct, iters, omega):
    x = og.zeros()
    sc = [og.zeros() for _ in range(9)]
    cnt, res = ctypes.c_int(), ctypes.c_double()
    st = og.L.orc_bicgstab(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.parr(sc), O.cint6(bct), O.cdbl6([0.0] * 6),
                           ctypes.c_double(omega), 1, iters, ctypes.c_double(0.0), ctypes.byref(cnt),
                           ctypes.byref(res))
    return x, cnt.value, res.value, st


P_BCT = [1, 0, 1, 1, 1, 1]  # X+ Dirichlet, other faces Neumann (the flow cases' pressure BC)


@pytest.mark.parametrize("name", list(TREES))
@pytest.mark.parametrize("omega", [1.0, 0.8])
def test_bicgstab_is_textbook_bicgstab(name, omega):
    og = OracleGrid(TREES[name])
    A, b = system(og, 3, P_BCT)
    op = Operator(og, A)
    bi = interior(b).ravel()
    dinv = 1.0 / interior(A[0]).ravel()
    for iters in (1, 4, 9):
        xo, cnt, res, st = oracle_bicgstab(og, A, b, P_BCT, iters, omega)
        xt, rt = textbook_bicgstab(op, dinv, bi, iters, omega)
        scale = np.max(np.abs(xt))
        assert st == 0 and cnt == iters + 1  # never converged at eps = 0: count = countMax + 1
        assert np.max(np.abs(interior(xo).ravel() - xt)) <= 1e-10 * scale, (name, iters)
        # IsConverged's residual is the recursive one, relative to |b|
        assert abs(res - np.linalg.norm(rt) / np.linalg.norm(bi)) <= 1e-9 * res
    # scipy's bicgstab is the same right-preconditioned form (a library routine)
    M = spla.LinearOperator(op.linop().shape, matvec=lambda v: omega * dinv * v, dtype=np.float64)
    xs, _ = spla.bicgstab(op.linop(), bi, x0=np.zeros_like(bi), M=M, maxiter=6, rtol=1e-300, atol=0.0)
    xo, _, _, _ = oracle_bicgstab(og, A, b, P_BCT, 6, omega)
    assert np.max(np.abs(interior(xo).ravel() - xs)) <= 1e-10 * np.max(np.abs(xs))


@pytest.mark.parametrize("mistake", [1, 2])
def test_bicgstab_pin_catches_planted_mistakes(plant, mistake):
    """1: beta with the wrong sign; 2: alpha and gamma swapped in the x update."""
    og = OracleGrid(TREES["two_level"])
    A, b = system(og, 3, P_BCT)
    op = Operator(og, A)
    xt, _ = textbook_bicgstab(op, 1.0 / interior(A[0]).ravel(), interior(b).ravel(), 5, 1.0)
    plant(mistake)
    xo, _, _, _ = oracle_bicgstab(og, A, b, P_BCT, 5, 1.0)
    assert np.max(np.abs(interior(xo).ravel() - xt)) > 1e-6 * np.max(np.abs(xt))


def oracle_jacobi(og, A, b, bct, iters, omega):
    x, x0 = og.zeros(), og.zeros()
    cnt, res = ctypes.c_int(), ctypes.c_double()
    og.L.orc_jacobi(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.ptr(x0), O.cint6(bct), O.cdbl6([0.0] * 6),
                    ctypes.c_double(omega), iters, ctypes.c_double(0.0), ctypes.byref(cnt), ctypes.byref(res))
    return x, cnt.value, res.value


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_jacobi_is_textbook_jacobi(name):
    """x_{k+1} = x_k + omega D^-1 (b - A x_k) on the Dirichlet system (all outer faces)."""
    og = OracleGrid(TREES[name])
    bct = [0] * 6
    A, b = system(og, 4, bct)
    op = Operator(og, A, (bct, [0.0] * 6))
    bi = interior(b).ravel()
    dinv = 1.0 / interior(A[0]).ravel()
    omega = 0.9
    for iters in (1, 6):
        xt = np.zeros_like(bi)
        for _ in range(iters):
            xt = xt + omega * dinv * (bi - op(xt))
        xo, cnt, res = oracle_jacobi(og, A, b, bct, iters, omega)
        assert cnt == iters + 1
        assert np.max(np.abs(interior(xo).ravel() - xt)) <= 1e-12 * np.max(np.abs(xt))
        want = np.linalg.norm(bi - op(xt)) / np.linalg.norm(bi)
        assert abs(res - want) <= 1e-10 * want


def test_jacobi_pin_catches_planted_mistake(plant):
    og = OracleGrid(TREES["flat"])
    bct = [0] * 6
    A, b = system(og, 4, bct)
    op = Operator(og, A, (bct, [0.0] * 6))
    bi = interior(b).ravel()
    dinv = 1.0 / interior(A[0]).ravel()
    xt = np.zeros_like(bi)
    for _ in range(4):
        xt = xt + 0.9 * dinv * (bi - op(xt))
    plant(3)
    xo, _, _ = oracle_jacobi(og, A, b, bct, 4, 0.9)
    assert np.max(np.abs(interior(xo).ravel() - xt)) > 1e-3 * np.max(np.abs(xt))


def laplacian(og, bct, bcv=None):
    """The 7-point Laplacian 1/dx^2 per block with the outer BC folded in (bca.f90)."""
    A = [og.zeros() for _ in range(7)]
    dx2 = (og.tb["dx"] ** 2)[:, None, None, None]
    for q in range(1, 7):
        A[q][:] = 1.0 / dx2
    A[0][:] = -6.0 / dx2
    b = og.zeros()
    og.L.orc_matrix_bc(og.ref(), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv or [0.0] * 6))
    return A, b


def oracle_fixed_average(og, A, b, iters, omega, eps=0.0):
    x = og.zeros()
    cnt, res = ctypes.c_int(), ctypes.c_double()
    og.L.orc_rbgs_fixed_average(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.cint6([1] * 6), O.cdbl6([0.0] * 6),
                                og.nb * N ** 3, ctypes.c_double(omega), iters, ctypes.c_double(eps),
                                ctypes.byref(cnt), ctypes.byref(res))
    return x, cnt.value, res.value


def textbook_rbsor_fixed_average(A, b, iters, omega):
    """Red-black SOR on one block (colour 0 = (i+j+k) odd, 1-based, first), then the mean removed,
    per iteration.  The outer BC is folded into A, so the ghosts never matter."""
    Ai = [a[0, G:G + N, G:G + N, G:G + N] for a in A]
    bi = b[0, G:G + N, G:G + N, G:G + N]
    u = np.zeros((N + 2,) * 3)
    k, j, i = np.meshgrid(*(np.arange(1, N + 1),) * 3, indexing="ij")
    masks = [((i + j + k) % 2 == 1), ((i + j + k) % 2 == 0)]
    for _ in range(iters):
        for m in masks:
            nb = (Ai[1] * u[1:-1, 1:-1, :-2] + Ai[2] * u[1:-1, 1:-1, 2:] + Ai[3] * u[1:-1, :-2, 1:-1] +
                  Ai[4] * u[1:-1, 2:, 1:-1] + Ai[5] * u[:-2, 1:-1, 1:-1] + Ai[6] * u[2:, 1:-1, 1:-1])
            inner = u[1:-1, 1:-1, 1:-1]
            new = inner - omega * (inner - (bi - nb) / Ai[0])
            inner[m] = new[m]
        u[1:-1, 1:-1, 1:-1] -= u[1:-1, 1:-1, 1:-1].mean()
    return u[1:-1, 1:-1, 1:-1]


def _neumann_one_block(seed):
    og = OracleGrid(ffvhc.Tree(root=(1, 1, 1), root_length=1.0))
    A, b = laplacian(og, [1] * 6)
    rng = np.random.default_rng(seed)
    b[INT] = rng.uniform(-1, 1, (1, N, N, N))
    b[INT] -= b[INT].mean()
    return og, A, b


def test_fixed_average_is_textbook_rbsor_with_mean_removal():
    og, A, b = _neumann_one_block(6)
    for iters in (1, 5):
        xo, cnt, _ = oracle_fixed_average(og, A, b, iters, 1.5)
        xt = textbook_rbsor_fixed_average(A, b, iters, 1.5)
        assert cnt == iters + 1
        assert abs(interior(xo).mean()) <= 1e-14 * np.max(np.abs(xt))
        assert np.max(np.abs(interior(xo)[0] - xt)) <= 1e-12 * np.max(np.abs(xt))


def test_fixed_average_pin_catches_planted_mistake(plant):
    og, A, b = _neumann_one_block(6)
    xt = textbook_rbsor_fixed_average(A, b, 5, 1.5)
    plant(4)
    xo, _, _ = oracle_fixed_average(og, A, b, 5, 1.5)
    assert np.max(np.abs(interior(xo)[0] - xt)) > 1e-6 * np.max(np.abs(xt))


def manufactured_error(tree):
    """Max error of the fixed-average RBGS solution of lap p = f with Neumann walls against
    p = cos(pi x) cos(pi y) cos(pi z) (zero normal derivative on the unit cube), both with their
    cell means removed."""
    from helpers import cell_centres
    og = OracleGrid(tree)
    A, b = laplacian(og, [1] * 6)
    x, y, z = cell_centres(og.tb)
    exact = np.cos(np.pi * x) * np.cos(np.pi * y) * np.cos(np.pi * z)
    vol = (og.tb["dx"] ** 3)[:, None, None, None]
    b[INT] = -3.0 * np.pi ** 2 * exact[INT]
    b[INT] -= np.sum(b[INT] * vol) / np.sum(vol * np.ones_like(b[INT]))
    xo, cnt, res = oracle_fixed_average(og, A, b, 20000, 1.3, 1e-11)
    assert res <= 1e-11, (cnt, res)
    e = interior(xo) - exact[INT]
    e -= e.mean()
    return np.max(np.abs(e))


@pytest.mark.slow
def test_fixed_average_manufactured_solution_converges_second_order():
    e16 = manufactured_error(ffvhc.Tree(root=(1, 1, 1), root_length=1.0))
    e32 = manufactured_error(ffvhc.Tree(root=(2, 2, 2), root_length=0.5))
    assert e32 < e16
    assert 3.5 < e16 / e32 < 4.5, (e16, e32)


# ------------------------------------------------------------------ Solver::Update invariants

FLOW_BC = {ffvhc.BC_UX: [(0, 1.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)],
           ffvhc.BC_UY: [(0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)],
           ffvhc.BC_UZ: [(0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)],
           ffvhc.BC_P: [(1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0), (1, 0.0), (1, 0.0)]}


def no_geometry(og):
    return (np.ones((6,) + og.shape), np.zeros((6,) + og.shape, np.int32), np.ones(og.shape, np.int32))


def arrays(og, ux, uy, uz, p):
    z = og.zeros
    a = {"ux": ux, "uy": uy, "uz": uz, "p": p, "b": z(), "lapp": z(), "t": z(),
         "uc": [z() for _ in range(3)], "ucp": [z() for _ in range(3)],
         "ud": [z() for _ in range(3)], "udp": [z() for _ in range(3)],
         "f": [z() for _ in range(3)], "v": [z() for _ in range(6)], "A": [z() for _ in range(7)],
         "cgs": [z() for _ in range(5)], "bis": [z() for _ in range(9)]}
    return a


def fill_initial(og, params, a):
    bl = lambda grp: ([params.bct[grp][f] for f in range(6)], [params.bcv[grp][f] for f in range(6)])
    for d, key in enumerate(("ux", "uy", "uz")):
        og.ghost_fill(a[key], 5, *bl(d))
    og.ghost_fill(a["p"], 1, *bl(ffvhc.BC_P))
    # face velocities of the initial field (what calc_ab_p of the previous step leaves)
    a["v"][0][:] = a["ux"]
    a["v"][1][:] = a["ux"]


def _uniform_flow(name, p0_scale, nsteps, solver=3, seed=8):
    og = OracleGrid(TREES[name])
    params = ffvhc.case_params(rho=1.0, mu=0.01, dt=0.002, bc=FLOW_BC)
    rng = np.random.default_rng(seed)
    a = arrays(og, np.ones(og.shape), og.zeros(), og.zeros(), p0_scale * rng.standard_normal(og.shape))
    fill_initial(og, params, a)
    orc = OracleStep(og, params, no_geometry(og), a, solver, 1.0, 400, 1e-12)
    infos = [orc.step(n) for n in range(nsteps)]
    return og, a, infos


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_update_keeps_uniform_flow_exactly(name):
    og, a, infos = _uniform_flow(name, 0.0, 2, solver=0)
    for key, want in (("ux", 1.0), ("uy", 0.0), ("uz", 0.0), ("p", 0.0), ("lapp", 0.0)):
        assert np.all(interior(a[key]) == want), key
    assert all(i.status == 0 for i in infos)


@pytest.mark.parametrize("name", ["flat", "two_level"])
def test_update_restores_uniform_flow_after_pressure_gradient(name):
    """The predictor subtracts grad p^n and bcut_remove_p adds it back, so a uniform flow with any
    initial pressure field is a uniform flow again after the projection (p -> the solution of a
    zero right-hand side)."""
    og, a, _ = _uniform_flow(name, 0.05, 1)
    assert np.max(np.abs(interior(a["ux"]) - 1.0)) < 1e-10
    assert np.max(np.abs(interior(a["uy"]))) < 1e-10
    assert np.max(np.abs(interior(a["uz"]))) < 1e-10


def test_uniform_flow_pin_catches_skipped_remove_p(plant):
    plant(6)
    og, a, _ = _uniform_flow("flat", 0.05, 1)
    assert np.max(np.abs(interior(a["ux"]) - 1.0)) > 1e-4


def _projection(name, solver, eps, seed=9, cuts=False):
    og = OracleGrid(TREES[name])
    params = ffvhc.case_params(rho=1.3, mu=0.01, dt=0.004, bc=FLOW_BC)
    rng = np.random.default_rng(seed)
    geo = geometry(og, rng) if cuts else no_geometry(og)
    a = arrays(og, 1.0 + 0.1 * rng.standard_normal(og.shape), 0.1 * rng.standard_normal(og.shape),
               0.1 * rng.standard_normal(og.shape), 0.05 * rng.standard_normal(og.shape))
    fill_initial(og, params, a)
    orc = OracleStep(og, params, geo, a, solver, 1.0, 3000, eps)
    info = orc.step(0)
    dx = og.tb["dx"][:, None, None, None]
    v = [interior(f) for f in a["v"]]
    div = (v[1] - v[0] + v[3] - v[2] + v[5] - v[4]) / dx / params.dt
    resid = interior(a["b"]) - stencil(a["A"], a["p"])
    return og, a, info, div, resid


@pytest.mark.parametrize("name,cuts", [("flat", False), ("two_level", False), ("flat", True), ("two_level", True)])
def test_projection_leaves_divergence_at_solver_tolerance(name, cuts):
    """After bcut_corr_u, div(v_face)/dt of every cell is the residual b - A p of the pressure
    solve (face gradients and matrix bands are the same r/dx^2 couplings, also at cut faces and
    solid cells), so its norm relative to |b| is the solver's residual."""
    og, a, info, div, resid = _projection(name, 3, 1e-11, cuts=cuts)
    bnorm = np.linalg.norm(interior(a["b"]))
    assert info.residualP <= 1e-11
    assert np.max(np.abs(div - resid)) <= 1e-10 * np.max(np.abs(interior(a["b"])))
    assert np.linalg.norm(div) / bnorm <= 2e-11


def test_projection_pin_catches_swapped_stages(plant):
    plant(5)
    og, a, info, div, resid = _projection("flat", 3, 1e-11)
    assert np.linalg.norm(div) / np.linalg.norm(interior(a["b"])) > 1e-3


def test_ghost_fill_pin_catches_swapped_order(plant):
    """Scalar3DUpdater5's C2F reads the coarse ghost F2C filled: with C2F first, a linear field is
    no longer reproduced at the coarse/fine faces."""
    from helpers import cell_centres
    og = OracleGrid(TREES["two_level"])
    x, y, z = cell_centres(og.tb)
    lin = 0.3 + 1.7 * x - 0.9 * y + 2.3 * z
    f = og.zeros()
    f[INT] = lin[INT]
    plant(8)
    og.ghost_fill(f, 5, [-1] * 6, [0.0] * 6)
    bad = 0
    for b in range(og.nb):
        for face in range(6):
            if og.tb["ftype"][b, face] != 2:
                continue
            ax = face // 2
            idx = [slice(G, G + N)] * 3
            idx[2 - ax] = (G + N) if face & 1 else (G - 1)
            bad += int(not np.allclose(f[(b,) + tuple(idx)], lin[(b,) + tuple(idx)], rtol=0, atol=1e-12))
    assert bad > 0
