"""Sphere front end (host): analytic cut positions, the bstl post-processing against the oracle's
restatement of bstl.f90 (itself pinned to the compiled reference), and the phase flood fill."""
import numpy as np
import pytest

import ffvhc
from oracle import oracle as O
from helpers import G, N, cell_centres, interior

NP = N + 2 * G


def test_cut_positions_lie_on_the_sphere():
    tree = ffvhc.Tree(root=(2, 2, 2), root_length=1.0)
    C, R = np.array([0.93, 1.04, 1.11]), 0.61
    c, cid, pid, counts = tree.geometry_sphere(C, R, cutoff=0.0, hole_filling=(0, 0, 0))
    tb = tree.table()
    x, y, z = cell_centres(tb)
    xyz = [x.reshape(tree.nblk, -1), y.reshape(tree.nblk, -1), z.reshape(tree.nblk, -1)]
    dx = tb["dx"][:, None]
    ncut = 0
    for f in range(6):
        ax, s = f // 2, (1.0 if f & 1 else -1.0)
        m = cid[f] != 0
        ncut += m.sum()
        p = [xyz[a] + (s * c[f] * dx if a == ax else 0.0) for a in range(3)]
        r = np.sqrt(sum((p[a] - C[a]) ** 2 for a in range(3)))
        # float32 positions in cell units: |r - R| <= float32 eps * dx
        assert np.all(np.abs(r[m] - R) <= 1e-6 * np.broadcast_to(dx, r.shape)[m])
        assert np.all(c[f][~m] == 1.0)
        assert np.all((c[f][m] >= 0.0) & (c[f][m] <= 1.0))
        assert np.all(c[f][m] == c[f][m].astype(np.float32))
    assert ncut > 0


def test_cut_faces_pair_up_across_the_surface():
    """A cut seen from a cell along +x at distance d is seen from the +x neighbour along -x at
    distance 1 - d (same surface point)."""
    tree = ffvhc.Tree(root=(1, 1, 1), root_length=1.0)
    c, cid, pid, _ = tree.geometry_sphere((0.5, 0.52, 0.47), 0.3, cutoff=0.0, hole_filling=(0, 0, 0))
    c = c.reshape(6, 1, NP, NP, NP)
    cid = cid.reshape(6, 1, NP, NP, NP)
    # +x (face 1) of (i) and -x (face 0) of (i+1), interior only
    a = interior(c[1])[..., :-1]
    b = interior(c[0])[..., 1:]
    ma = interior(cid[1])[..., :-1] != 0
    mb = interior(cid[0])[..., 1:] != 0
    both = ma & mb
    assert both.sum() > 0
    assert np.allclose(a[both] + b[both], 1.0, atol=2e-7)


def test_phase_is_inside_outside():
    tree = ffvhc.Tree(root=(2, 2, 2), root_length=1.0, regions=[dict(kind="ball", a=(1.0, 1.0, 1.0), r=0.7, level=1)])
    C, R = (1.0, 1.0, 1.0), 0.37
    c, cid, pid, counts = tree.geometry_sphere(C, R, cutoff=0.01, hole_filling=(1, 1, 0), seed=(0.01, 0.01, 0.01))
    x, y, z = cell_centres(tree.table())
    inside = (x - C[0]) ** 2 + (y - C[1]) ** 2 + (z - C[2]) ** 2 < R * R
    got = interior(pid.reshape(tree.nblk, NP, NP, NP))
    want = np.where(interior(inside), 0, 1)
    assert np.array_equal(got, want)
    assert counts[0] == want.sum() and counts[1] == want.size - want.sum()


def _oracle_postprocess(c, cid, eps, v1, v2):
    """bstl_cutoff, detect_zerocut, fill_holes_v1 and iterated v2 on one block with Neumann-0
    ghosts (a single-block tree), through the oracle."""
    L = O.lib()
    sz = (3 * __import__("ctypes").c_int)(N, N, N)
    cl = [np.ascontiguousarray(c[f, 0]) for f in range(6)]
    il = [np.ascontiguousarray(cid[f, 0]) for f in range(6)]

    def exch():
        for arrs in (cl, il):
            for a in arrs:
                v = a.reshape(NP, NP, NP)
                for l in range(1, G + 1):
                    v[G - l, :, :] = v[G, :, :]
                    v[G + N - 1 + l, :, :] = v[G + N - 1, :, :]
                    v[:, G - l, :] = v[:, G, :]
                    v[:, G + N - 1 + l, :] = v[:, G + N - 1, :]
                    v[:, :, G - l] = v[:, :, G]
                    v[:, :, G + N - 1 + l] = v[:, :, G + N - 1]

    exch()
    if eps > 0:
        L.orc_bstl_cutoff(O.parr(cl), O.parr(il, O.ctypes.c_int), O.ctypes.c_double(eps), sz, G)
        exch()
    L.orc_bstl_detect_zerocut(O.parr(cl), O.parr(il, O.ctypes.c_int), sz, G)
    exch()
    n1 = L.orc_bstl_fill_holes_v1(O.parr(cl), O.parr(il, O.ctypes.c_int), int(v1), sz, G)
    exch()
    n2 = 0
    for _ in range(1000):
        k = L.orc_bstl_fill_holes_v2(O.parr(cl), O.parr(il, O.ctypes.c_int), int(v2), sz, G)
        exch()
        n2 += k
        if k == 0 or not v2:
            break
    return np.stack(cl), np.stack(il), n1, n2


@pytest.mark.parametrize("radius,center", [(0.31, (0.5, 0.5, 0.5)), (0.2, (0.43, 0.61, 0.5))])
def test_bstl_postprocessing_matches_oracle(radius, center):
    tree = ffvhc.Tree(root=(1, 1, 1), root_length=1.0)
    raw_c, raw_id, _, _ = tree.geometry_sphere(center, radius, cutoff=0.0, hole_filling=(0, 0, 0))
    c, cid, _, counts = tree.geometry_sphere(center, radius, cutoff=0.05, hole_filling=(1, 1, 0))
    oc, oid, n1, n2 = _oracle_postprocess(raw_c, raw_id, 0.05, 1, 1)
    live = np.zeros((NP, NP, NP), bool)
    live[G:G + N, G:G + N, G:G + N] = True
    live = live.ravel()
    assert np.array_equal(cid[:, 0][:, live], oid[:, live])
    assert np.array_equal(c[:, 0][:, live], oc[:, live])
    assert counts[2] == n1 and counts[3] == n2


def test_cut_array_ghosts_follow_the_level_updaters():
    """The cut arrays' ghost cells after the front end are what the reference's ImposeBoundaryCondition
    leaves there, on a tree whose sphere surface crosses a level interface: cut distances as the
    oracle's Updater1 fill of the interior (F2C means, same level, C2F trilinear, Neumann 0 outside),
    cut ids the same values in double truncated to int (the Scalar3DUpdater1 template on
    LocalScalar3D<int>, Scalar3DUpdater1.h:164-252, 503-700).  Face ghosts of both layers; edges and
    corners are not updated by the reference either."""
    tree = ffvhc.Tree(root=(2, 2, 2), root_length=1.0, regions=[dict(kind="ball", a=(0.6, 0.6, 0.6), r=0.35,
                                                                      level=1)])
    assert tree.max_level == 1
    c, cid, pid, counts = tree.geometry_sphere((0.8, 0.85, 0.75), 0.47, cutoff=0.0, hole_filling=(0, 0, 0))
    tb = tree.table()
    tb["g"] = G
    tab = dict(tb)
    tab.update(ftype=tb["ftype"].ravel(), nbr=tb["nbr"].ravel(), sub=tb["sub"].ravel(), org=tb["org"].ravel())
    grid = O.make_grid(tab)
    L = O.lib()
    shp = (tree.nblk, NP, NP, NP)
    idx = np.arange(NP) - G
    inside = (idx >= 0) & (idx < N)
    nout = (~inside)[:, None, None].astype(int) + (~inside)[None, :, None] + (~inside)[None, None, :]
    face_ghost = nout == 1
    interior_m = nout == 0
    crossing = 0
    for f in range(6):
        cf = c[f].reshape(shp)
        want = np.zeros(shp)
        want[:, interior_m] = cf[:, interior_m]
        L.orc_ghost_fill(ctypes_ref(grid), O.ptr(want), 1, O.cint6([1] * 6), O.cdbl6([0.0] * 6))
        assert np.array_equal(cf[:, face_ghost], want[:, face_ghost]), f
        idf = cid[f].reshape(shp)
        wi = np.zeros(shp)
        wi[:, interior_m] = idf[:, interior_m]
        L.orc_ghost_fill(ctypes_ref(grid), O.ptr(wi), 1, O.cint6([1] * 6), O.cdbl6([0.0] * 6))
        assert np.array_equal(idf[:, face_ghost], np.trunc(wi[:, face_ghost]).astype(np.int32)), f
        # the surface reaches ghosts at level interfaces: interpolated (not float32) cut values there
        for b in range(tree.nblk):
            for ff in range(6):
                if tb["ftype"][b, ff] in (2, 3):
                    ax = 2 - (ff >> 1)
                    sl = [slice(G, G + N)] * 3
                    sl[ax] = slice(G + N, NP) if ff & 1 else slice(0, G)
                    v = cf[(b,) + tuple(sl)]
                    crossing += int(np.count_nonzero((v != 1.0) & (v != v.astype(np.float32))))
    assert crossing > 0


def ctypes_ref(grid):
    import ctypes
    return ctypes.byref(grid)
