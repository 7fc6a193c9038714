"""Grid-level oracle pins that do not depend on the reference's (unbuildable) C++ driver.

* Ghost fill: same-level copies, F2C means and C2F prolongation reproduce a global linear field
  exactly at every inner-face ghost (Scalar3DUpdater1: trilinear; Scalar3DUpdater5: gradient
  step).  A wrong subface, axis or layer index puts a value at the wrong place and fails.
* RBGS on one block with omega = 1 equals textbook red-black Gauss-Seidel (numpy, vectorised).
* Jacobi-PCG with PreCount = 1, omega = 1 reproduces scipy.sparse.linalg.cg with a Jacobi
  preconditioner iterate by iterate (a library routine, not the reference).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import ffvhc
from oracle import oracle as O
from helpers import G, N, OracleGrid, cell_centres, interior


def _trees():
    return {
        "flat": ffvhc.Tree(root=(3, 2, 2), root_length=1.0),
        "two_level": ffvhc.Tree(root=(3, 3, 2), root_length=1.0,
                                regions=[dict(kind="ball", a=(1.4, 1.6, 1.0), r=0.45, level=1)]),
        "three_level": ffvhc.Tree(root=(2, 2, 2), root_length=1.0,
                                  regions=[dict(kind="ball", a=(0.9, 1.1, 1.05), r=0.2, level=2)]),
    }


TREES = _trees()


@pytest.mark.parametrize("name", list(TREES))
@pytest.mark.parametrize("updater", [1, 5])
def test_ghost_fill_reproduces_linear_fields(name, updater):
    og = OracleGrid(TREES[name])
    x, y, z = cell_centres(og.tb)
    lin = 0.3 + 1.7 * x - 0.9 * y + 2.3 * z
    f = og.zeros()
    sl = (slice(None),) + (slice(G, G + N),) * 3
    f[sl] = lin[sl]
    og.ghost_fill(f, updater, [-1] * 6, [0.0] * 6)
    tb = og.tb
    for b in range(og.nb):
        for face in range(6):
            if tb["ftype"][b, face] == 0:
                continue
            ax = face // 2
            layer = (G + N) if face & 1 else (G - 1)
            idx = [slice(G, G + N)] * 3
            idx[2 - ax] = layer  # arrays are (k, j, i)
            got = f[(b,) + tuple(idx)]
            want = lin[(b,) + tuple(idx)]
            assert np.allclose(got, want, rtol=0, atol=1e-12), (name, b, face, tb["ftype"][b, face])


def test_rbgs_one_block_is_textbook_red_black_gs():
    t = ffvhc.Tree(root=(1, 1, 1), root_length=1.0)
    og = OracleGrid(t)
    rng = np.random.default_rng(5)
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = 1.0
    A[0][:] = -6.0
    b = rng.uniform(-1, 1, og.shape)
    x = og.zeros()
    L = O.lib()
    cnt = ctypes.c_int()
    res = ctypes.c_double()
    # Dirichlet 0 ghosts: x(0) = -x(1) etc., folded matrix not applied (ghosts carry the BC)
    L.orc_rbgs(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), O.cint6([0] * 6), O.cdbl6([0.0] * 6), ctypes.c_double(1.0),
               25, ctypes.c_double(0.0), ctypes.byref(cnt), ctypes.byref(res))
    # textbook red-black GS, u on a 16^3 box with antisymmetric (Dirichlet 0 at the face) ghosts
    u = np.zeros((N + 2,) * 3)
    bb = interior(b, g=G)[0]
    kk, jj, ii = np.meshgrid(np.arange(1, N + 1), np.arange(1, N + 1), np.arange(1, N + 1), indexing="ij")
    masks = [((ii + jj + kk) % 2 == 1), ((ii + jj + kk) % 2 == 0)]  # colour 0 first (bils.f90:88)
    for _ in range(25):
        for m in masks:
            u[0, :, :], u[-1, :, :] = -u[1, :, :], -u[-2, :, :]
            u[:, 0, :], u[:, -1, :] = -u[:, 1, :], -u[:, -2, :]
            u[:, :, 0], u[:, :, -1] = -u[:, :, 1], -u[:, :, -2]
            nb_sum = (u[:-2, 1:-1, 1:-1] + u[2:, 1:-1, 1:-1] + u[1:-1, :-2, 1:-1] + u[1:-1, 2:, 1:-1] +
                      u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 2:])
            new = (bb - nb_sum) / -6.0
            inner = u[1:-1, 1:-1, 1:-1]
            inner[m] = new[m]
    assert cnt.value == 26  # never converges at eps = 0: count = countMax + 1
    assert np.allclose(interior(x)[0], u[1:-1, 1:-1, 1:-1], rtol=1e-13, atol=1e-13)


def _global_matrix(og, A):
    """Sparse matrix of the 7-band fields over all interior cells (single-level grid)."""
    nb, tb = og.nb, og.tb
    ncell = nb * N ** 3
    gid = np.arange(ncell).reshape(nb, N, N, N)
    rows, cols, vals = [], [], []
    Ai = [interior(a) for a in A]
    rows.append(gid.ravel())
    cols.append(gid.ravel())
    vals.append(Ai[0].ravel())
    offs = {1: (0, 0, -1), 2: (0, 0, 1), 3: (0, -1, 0), 4: (0, 1, 0), 5: (-1, 0, 0), 6: (1, 0, 0)}
    for q, (dk, dj, di) in offs.items():
        face = q - 1
        for b in range(nb):
            for k in range(N):
                for j in range(N):
                    for i in range(N):
                        v = Ai[q][b, k, j, i]
                        if v == 0.0:
                            continue
                        k2, j2, i2, b2 = k + dk, j + dj, i + di, b
                        if not (0 <= k2 < N and 0 <= j2 < N and 0 <= i2 < N):
                            assert tb["ftype"][b, face] == 1
                            b2 = tb["nbr"][b, face, 0]
                            k2, j2, i2 = k2 % N, j2 % N, i2 % N
                        rows.append([gid[b, k, j, i]])
                        cols.append([gid[b2, k2, j2, i2]])
                        vals.append([v])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(ncell, ncell))


def test_jacobi_pcg_matches_scipy_cg():
    t = ffvhc.Tree(root=(2, 1, 1), root_length=1.0)
    og = OracleGrid(t)
    rng = np.random.default_rng(11)
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = 1.0
    A[0][:] = -6.0 - rng.uniform(0.0, 0.3, og.shape)  # SPD-shift keeps -A definite
    b = og.zeros()
    b[(slice(None),) + (slice(G, G + N),) * 3] = rng.uniform(-1, 1, (og.nb, N, N, N))
    bct, bcv = [0] * 6, [0.0] * 6
    og.L.orc_matrix_bc(og.ref(), O.parr(A), O.ptr(b), O.cint6(bct), O.cdbl6(bcv))
    M = _global_matrix(og, A)
    rhs = interior(b).ravel()
    scratch = [og.zeros() for _ in range(5)]
    for iters in (5, 12):
        x = og.zeros()
        cnt, res = ctypes.c_int(), ctypes.c_double()
        og.L.orc_cg(og.ref(), O.ptr(x), O.parr(A), O.ptr(b), *[O.ptr(s) for s in scratch], O.cint6(bct), O.cdbl6(bcv),
                    ctypes.c_double(1.0), 1, iters, ctypes.c_double(1e-30), ctypes.byref(cnt), ctypes.byref(res))
        # scipy on -A (SPD) with the Jacobi preconditioner, same iteration count
        Minv = sp.diags(1.0 / (-M.diagonal()))
        xs, _ = spla.cg(-M, -rhs, x0=np.zeros_like(rhs), M=Minv, maxiter=iters, rtol=1e-300, atol=0.0)
        assert np.allclose(interior(x).ravel(), xs, rtol=1e-9, atol=1e-11)
