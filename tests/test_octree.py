"""Block octree (host code of the product): neighbour tables, 2:1 balance, ordering.

BCMTools, whose octree the reference uses, is not vendored in the reference, so the tables are
pinned here by invariants and a brute-force geometric neighbour search instead of reference
output: same-level neighbours share a full face, finer neighbours sit in the face quarter named
by their subface (origins along (y,z) of X faces, (z,x) of Y faces, (x,y) of Z faces,
Scalar3DUpdater1.cpp:264-343), coarse/fine links are symmetric, and Hilbert order walks
face-adjacent blocks.
"""
import numpy as np
import pytest

import ffvhc

TANG = {0: (1, 2), 1: (2, 0), 2: (0, 1)}  # face axis -> (origin0 axis, origin1 axis)


def boxes(tb):
    size = tb["dx"] * tb["n"]
    lo = tb["org"]
    return lo, lo + size[:, None], size


def brute_neighbours(tb, b, f, periodic=(0, 0, 0), extent=None):
    """All blocks touching face f of block b (by geometry), with their overlap boxes."""
    lo, hi, size = boxes(tb)
    ax = f // 2
    plane = hi[b, ax] if f & 1 else lo[b, ax]
    other = lo[:, ax] if f & 1 else hi[:, ax]
    t = [d for d in range(3) if d != ax]
    eps = 1e-9
    hit = np.abs(other - plane) < eps
    for d in t:
        hit &= (lo[:, d] < hi[b, d] - eps) & (hi[:, d] > lo[b, d] + eps)
    return np.flatnonzero(hit)


@pytest.mark.parametrize("kind", ["flat", "ball", "shell", "boxes"])
def test_neighbour_tables_match_brute_force(kind):
    if kind == "flat":
        t = ffvhc.Tree(root=(4, 3, 2), root_length=2.0)
    elif kind == "ball":
        t = ffvhc.Tree(root=(4, 4, 4), root_length=1.0, regions=[dict(kind="ball", a=(2, 2, 2), r=0.9, level=2)])
    elif kind == "shell":
        t = ffvhc.Tree(root=(4, 2, 2), root_length=1.0, regions=[dict(kind="shell", a=(1.3, 1, 1), r=0.5, dr=0.1,
                                                                     level=2)])
    else:
        t = ffvhc.Tree(root=(4, 2, 2), root_length=1.0, min_level=0,
                       regions=[dict(kind="box", a=(0.5, 0.5, 0.5), b=(3.0, 1.5, 1.5), level=1),
                                dict(kind="ball", a=(1, 1, 1), r=0.3, level=2)])
    tb = t.table()
    lo, hi, size = boxes(tb)
    nb = tb["nblk"]
    for b in range(nb):
        for f in range(6):
            ft = tb["ftype"][b, f]
            geo = brute_neighbours(tb, b, f)
            if ft == 0:
                assert geo.size == 0
            elif ft == 1:
                assert list(geo) == [tb["nbr"][b, f, 0]]
                assert tb["level"][geo[0]] == tb["level"][b]
            elif ft == 2:
                c = tb["nbr"][b, f, 0]
                assert list(geo) == [c] and tb["level"][c] == tb["level"][b] - 1
                # this block is quarter `sub` of the coarse face
                s = tb["sub"][b, f]
                t0, t1 = TANG[f // 2]
                o0 = int(round((lo[b, t0] - lo[c, t0]) / size[b]))
                o1 = int(round((lo[b, t1] - lo[c, t1]) / size[b]))
                assert s == o0 + 2 * o1
                assert tb["ftype"][c, f ^ 1] == 3 and tb["nbr"][c, f ^ 1, s] == b
            else:
                fine = tb["nbr"][b, f]
                assert sorted(geo) == sorted(fine)
                t0, t1 = TANG[f // 2]
                for s in range(4):
                    q = fine[s]
                    assert tb["level"][q] == tb["level"][b] + 1
                    assert int(round((lo[q, t0] - lo[b, t0]) / size[q])) == (s & 1)
                    assert int(round((lo[q, t1] - lo[b, t1]) / size[q])) == (s >> 1)
                    assert tb["ftype"][q, f ^ 1] == 2 and tb["sub"][q, f ^ 1] == s


def test_two_to_one_balance_and_cover():
    t = ffvhc.Tree(root=(2, 2, 2), root_length=1.0, regions=[dict(kind="ball", a=(0.3, 0.3, 0.3), r=0.05,
                                                                 level=4)])
    tb = t.table()
    lo, hi, size = boxes(tb)
    # leaves tile the domain exactly
    assert np.isclose(np.sum(size ** 3), 8.0)
    for b in range(tb["nblk"]):
        for f in range(6):
            for q in brute_neighbours(tb, b, f):
                assert abs(int(tb["level"][q]) - int(tb["level"][b])) <= 1


def test_c2_block_counts():
    """Standalone-Poisson grid: roots 8^3, level 1 inside a ball of radius L/4 at the centre."""
    t0 = ffvhc.Tree(root=(8, 8, 8), root_length=1.0, regions=[dict(kind="ball", a=(4, 4, 4), r=2.0, level=1)])
    assert t0.nblk == 960
    t1 = ffvhc.Tree(root=(8, 8, 8), root_length=1.0, margin=1.0,
                    regions=[dict(kind="ball", a=(4, 4, 4), r=2.0, level=1)])
    assert t1.nblk == 1128


def test_hilbert_order_is_face_continuous():
    t = ffvhc.Tree(root=(1, 1, 1), root_length=1.0, min_level=3, ordering=1)
    tb = t.table()
    lo = tb["org"]
    d = np.abs(np.diff(lo, axis=0)).sum(axis=1) / (tb["dx"][0] * 16)
    assert np.allclose(d, 1.0)
    # every cell visited once
    assert len({tuple(np.round(p, 9)) for p in lo}) == 512


def test_morton_order():
    t = ffvhc.Tree(root=(1, 1, 1), root_length=1.0, min_level=2, ordering=0)
    tb = t.table()
    ijk = np.round(tb["org"] / (tb["dx"][0] * 16)).astype(int)
    key = [sum(((i >> b) & 1) << (3 * b) | ((j >> b) & 1) << (3 * b + 1) | ((k >> b) & 1) << (3 * b + 2)
               for b in range(2)) for i, j, k in ijk]
    assert key == sorted(key) and len(set(key)) == 64


def test_roots_in_x_fastest_order_and_periodic_wrap():
    t = ffvhc.Tree(root=(3, 2, 1), root_length=1.0, periodic=(1, 0, 0))
    tb = t.table()
    assert np.allclose(tb["org"][:, 0], [0, 1, 2, 0, 1, 2])
    # periodic in x: X_M of the first root is the last root of the row
    assert tb["ftype"][0, 0] == 1 and tb["nbr"][0, 0, 0] == 2
    assert tb["ftype"][0, 2] == 0


def test_bad_spec_is_rejected():
    with pytest.raises(ffvhc.FFVError):
        ffvhc.Tree(root=(0, 1, 1))
