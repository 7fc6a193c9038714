"""Flow-past-a-sphere cases of the reference's example (sphere.tpp: uniform inflow on X-, outflow
on X+, slip-free Neumann elsewhere) on block octrees built by the library.

* single_level_sphere: RootBlock 8x4x4 of length 2D, 16^3 cells per block (h = D/8), sphere of
  diameter D at (4D, 4D, 4D) -- the reference's single-level flow past a sphere.
* three_level_sphere: roots of length D refined toward the sphere and its wake in the manner of
  SphereDivider3 (SphereDivider3.h:72-112): blocks cut by the sphere surface go to the finest
  level (sphericalBoxList[0]), blocks meeting a surrounding ball or a wake box go to that
  region's level (sphericalBoxList[1..], boundingBoxList), 2:1 balanced.  At full size (roots
  16x8x8, D = 1) this is 4,048 blocks = 16.6 M cells, h = D/64 at the sphere.  Blocks entirely
  inside the sphere would be LEAF_NO_ACTIVE in the reference; at these block sizes none is.
"""
import ffvhc

FLOW_BC = {ffvhc.BC_UX: [(0, 1.0)] + [(1, 0.0)] * 5,
           ffvhc.BC_UY: [(0, 0.0)] + [(1, 0.0)] * 5,
           ffvhc.BC_UZ: [(0, 0.0)] + [(1, 0.0)] * 5,
           ffvhc.BC_P: [(1, 0.0), (0, 0.0)] + [(1, 0.0)] * 4}


def flow_params(re, dt, u0=1.0, d=1.0, rho=1.0):
    """Case parameters of a flow at Reynolds number re = rho u0 d / mu."""
    return ffvhc.case_params(rho=rho, mu=rho * u0 * d / re, dt=dt, bc=FLOW_BC)


def single_level_sphere(radius=0.5):
    tree = ffvhc.Tree(root=(8, 4, 4), root_length=2.0)
    c, cid, pid, counts = tree.geometry_sphere(center=(4.0, 4.0, 4.0), radius=radius, cut_id=1)
    return tree, (c, cid, pid), counts


def sphere_regions(center, d, length, levels):
    """SphereDivider3-style regions for a tree of `levels` levels (1-3): three_level_regions, with
    the finest level capped at levels - 1 (two levels: the sphere surface, the ball and the wake box
    at level 1)."""
    if levels <= 1:
        return []
    top = levels - 1
    out = []
    for r in three_level_regions(center, d, length):
        if r["level"] <= top:
            out.append(r)
        elif r["kind"] == "shell":
            out.append(dict(r, level=top))
    return out


def three_level_regions(center, d=1.0, length=16.0):
    """SphereDivider3-style refinement: the sphere surface at level 2, a ball of radius 2d at
    level 1, the wake box x in [cx, cx + 12d], |y - cy|, |z - cz| <= 2d at level 1 and the near
    wake x in [cx - d/2, cx + 6d], |y - cy|, |z - cz| <= d at level 2 (x clipped to the domain)."""
    cx, cy, cz = center
    return [dict(kind="shell", a=center, r=0.5 * d, dr=0.0, level=2),
            dict(kind="ball", a=center, r=2.0 * d, level=1),
            dict(kind="box", a=(cx, cy - 2 * d, cz - 2 * d), b=(min(cx + 12 * d, length), cy + 2 * d, cz + 2 * d),
                 level=1),
            dict(kind="box", a=(cx - 0.5 * d, cy - d, cz - d), b=(min(cx + 6 * d, length), cy + d, cz + d),
                 level=2)]


def three_level_sphere(root=(16, 8, 8), root_length=1.0, d=1.0, geometry=True):
    """The three-level sphere + wake tree on roots of edge root_length (domain root * root_length),
    sphere of diameter d at (X / 4, Y / 2, Z / 2) of the domain, refinement regions scaled with d.
    Returns (tree, geometry or None, counts or None)."""
    center = (0.25 * root[0] * root_length, 0.5 * root[1] * root_length, 0.5 * root[2] * root_length)
    tree = ffvhc.Tree(root=root, root_length=root_length,
                      regions=three_level_regions(center, d, root[0] * root_length))
    if not geometry:
        return tree, None, None
    c, cid, pid, counts = tree.geometry_sphere(center=center, radius=0.5 * d, cut_id=1)
    return tree, (c, cid, pid), counts


def cfl_dt(tree, cfl, u0=1.0):
    """Time step for a CFL number on the finest cells: cfl * h_min / u0."""
    return cfl * float(tree.table()["dx"].min()) / u0


def served_case(target_blocks, levels, d=1.0):
    """A flow past a sphere of diameter d with about target_blocks blocks on `levels` levels, for
    the served stream: roots (2k, k, k) with the sphere at (X / 4, Y / 2, Z / 2); one level: root
    edge 2d (h = d / 8, the reference's single-level case); more levels: root edge d, 2d or 4d with
    the sphere + wake refinement of sphere_regions.  (k, root edge) is the pair whose block count is
    closest to the target (in log scale).  Returns (tree, center)."""
    import math
    best = None
    for rl in ((2.0 * d,) if levels <= 1 else (d, 2.0 * d, 4.0 * d)):
        for k in range(2, 64):
            root = (2 * k, k, k)
            center = (0.25 * root[0] * rl, 0.5 * root[1] * rl, 0.5 * root[2] * rl)
            tree = ffvhc.Tree(root=root, root_length=rl, regions=sphere_regions(center, d, root[0] * rl, levels))
            if tree.max_level + 1 != levels:
                continue
            err = abs(math.log(tree.nblk / float(target_blocks)))
            if best is None or err < best[0]:
                best = (err, tree, center)
            if tree.nblk > 2 * target_blocks:
                break
    return best[1], best[2]
