"""Phase flood fill on the device (FillCellFlag / InitPhase, Solver.cpp:1182-1287, 1404-1537).

* Against a breadth-first search written here from the rules (pure Python, small trees): from the
  seed cell, a move to the next cell along an axis is taken unless the entered cell's face toward
  the cell it comes from carries the boundary id; across a block face the entered cell is the one
  holding the point one cell over (same level, fine to coarse), or each of the 2 x 2 fine cells
  facing a coarse cell (coarse to fine).  Seeded cut ids close off some regions (walled cells
  and a walled box).
* Against the host geometry front end's own search on the flow-past-a-sphere geometries (one and
  three levels): the same fluid cells, bit for bit."""
from collections import deque

import numpy as np
import pytest

import ffvhc
from helpers import G, N, OracleGrid

pytestmark = pytest.mark.gpu


def bfs_phase(tb, cid, target, seed):
    """Reference search on padded cut ids cid[f][blk] (ghost width G); returns fluid[blk, k, j, i]."""
    nb = tb["nblk"]
    org, dx, ftype, nbr = tb["org"], tb["dx"], tb["ftype"], tb["nbr"]
    fluid = np.zeros((nb, N, N, N), bool)
    interior = cid[:, :, G:G + N, G:G + N, G:G + N]  # [f][blk][k][j][i]

    def blocked(b, i, j, k, f):
        return interior[f, b, k, j, i] == target

    start = None
    for b in range(nb):
        q = [int(np.floor((seed[a] - org[b, a]) / dx[b])) for a in range(3)]
        if all(0 <= v < N for v in q):
            start = (b, q[0], q[1], q[2])
            break
    assert start is not None
    fluid[start[0], start[3], start[2], start[1]] = True
    dq = deque([start])

    def visit(b, i, j, k, f_from):
        if not fluid[b, k, j, i] and not blocked(b, i, j, k, f_from ^ 1):
            fluid[b, k, j, i] = True
            dq.append((b, i, j, k))

    while dq:
        b, i, j, k = dq.popleft()
        cur = (i, j, k)
        for f in range(6):
            ax = f >> 1
            p = list(cur)
            p[ax] += 1 if f & 1 else -1
            if 0 <= p[ax] < N:
                visit(b, p[0], p[1], p[2], f)
                continue
            ft = ftype[b, f]
            if ft == 0:
                continue
            x = [org[b, a] + (cur[a] + 0.5) * dx[b] for a in range(3)]
            x[ax] += (1.0 if f & 1 else -1.0) * dx[b]
            if ft in (1, 2):  # same level, coarser
                nbk = nbr[b, f, 0]
                q = [min(max(int(np.floor((x[a] - org[nbk, a]) / dx[nbk])), 0), N - 1) for a in range(3)]
                visit(nbk, q[0], q[1], q[2], f)
            else:
                for s in range(4):
                    nbk = nbr[b, f, s]
                    for d1 in (0, 1):
                        for d2 in (0, 1):
                            y = list(x)
                            o = 0
                            for a in range(3):
                                if a == ax:
                                    y[a] = org[b, a] + (cur[a] + 0.5) * dx[b] + (0.75 if f & 1 else -0.75) * dx[b]
                                else:
                                    y[a] += (0.25 if (d1 if o == 0 else d2) else -0.25) * dx[b]
                                    o += 1
                            q = [int(np.floor((y[a] - org[nbk, a]) / dx[nbk])) for a in range(3)]
                            if all(0 <= v < N for v in q):
                                visit(nbk, q[0], q[1], q[2], f)
    return fluid


def closed_off_ids(og, seed):
    """Cut ids that close regions off: ids 1 and 2 on a quarter of the faces at random, every face
    of 5% of the cells id 1 (cells no move enters), and id 1 on the faces of the cells of a box that
    look out of the box (no move enters the box from outside)."""
    rng = np.random.default_rng(seed)
    cid = np.zeros((6,) + og.shape, np.int32)
    for f in range(6):
        m = rng.random(og.shape) < 0.25
        cid[f][m] = rng.integers(1, 3, int(m.sum()))
    walled = rng.random(og.shape) < 0.05
    cid[:, walled] = 1
    b = int(rng.integers(0, og.nb))
    lo, hi = G + 4, G + 10  # box of cells [4, 10)^3 of block b
    for f in range(6):
        ax = 2 - (f >> 1)  # array axis of face f's normal (k, j, i order)
        sl = [slice(lo, hi)] * 3
        sl[ax] = slice(hi - 1, hi) if f & 1 else slice(lo, lo + 1)
        cid[(f, b) + tuple(sl)] = 1
    return cid


@pytest.fixture(scope="module")
def ctx():
    c = ffvhc.Context(max_blocks=8192)
    yield c
    c.close()


@pytest.mark.parametrize("name,seed", [("flat", 3), ("two_level", 4), ("two_level", 5)])
def test_phase_fill_matches_search(ctx, name, seed):
    tree = {"flat": ffvhc.Tree(root=(3, 2, 1), root_length=1.0),
            "two_level": ffvhc.Tree(root=(2, 2, 1), root_length=1.0,
                                    regions=[dict(kind="ball", a=(1.0, 1.0, 0.5), r=0.3, level=1)])}[name]
    og = OracleGrid(tree)
    cid = closed_off_ids(og, seed)
    c = np.ones((6,) + og.shape)
    c[cid != 0] = 0.5
    case = ctx.add_case(tree, ffvhc.case_params(), geometry=(c, cid, np.ones(og.shape, np.int32)))
    try:
        sd = (0.01, 0.01, 0.01)
        counts = ctx.fill_phase(case, sd, cut_id=1)
        got = ctx.download_phase(case)[:, G:G + N, G:G + N, G:G + N].astype(bool)
        want = bfs_phase(og.tb, cid, 1, sd)
        assert 0 < want.sum() < want.size  # some cells are closed off
        assert np.array_equal(got, want)
        assert counts[0] == want.sum() and counts[0] + counts[1] == want.size
    finally:
        ctx.remove_case(case)


@pytest.mark.parametrize("levels", [1, 3])
def test_phase_fill_matches_host_geometry(ctx, levels):
    import flows
    if levels == 1:
        tree, center = flows.served_case(128, 1)
    else:
        tree, center = flows.served_case(400, 3)
    assert tree.max_level + 1 == levels
    c, cid, pid, counts = tree.geometry_sphere(center=center, radius=0.5, cut_id=1)
    case = ctx.add_case(tree, ffvhc.case_params(), geometry=(c, cid, None))
    try:
        seed = tuple(float(v) + 1e-9 for v in tree.spec.origin)
        dc = ctx.fill_phase(case, seed, cut_id=1)
        got = ctx.download_phase(case)
        shp = got.shape
        want = pid.reshape(shp)
        inner = (slice(None),) + (slice(G, G + N),) * 3
        assert np.array_equal(got[inner], want[inner])
        assert dc[0] == counts[0] and dc[1] == counts[1] and counts[1] > 0
    finally:
        ctx.remove_case(case)
