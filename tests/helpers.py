"""Shared helpers for the tests: grids, padded fields, oracle grid handles."""
import ctypes

import numpy as np

import ffvhc
from oracle import oracle as O

G = 2  # ghost layers of the reference layout (NumberOfVirtualCells, FFVConfig.cpp:44)
N = 16


def cell_centres(tb, g=G):
    """x, y, z of every padded cell, shape (nb, n+2g, n+2g, n+2g) in (k, j, i) order."""
    n = tb["n"]
    idx = np.arange(-g, n + g) + 0.5
    dx = tb["dx"][:, None, None, None]
    org = tb["org"]
    z = org[:, 2, None, None, None] + dx * idx[None, :, None, None]
    y = org[:, 1, None, None, None] + dx * idx[None, None, :, None]
    x = org[:, 0, None, None, None] + dx * idx[None, None, None, :]
    shp = (tb["nblk"], n + 2 * g, n + 2 * g, n + 2 * g)
    return np.broadcast_to(x, shp), np.broadcast_to(y, shp), np.broadcast_to(z, shp)


def live_mask(n=N, g=G):
    """Interior cells plus the first ghost layer of each face (what the device stores)."""
    m = np.zeros((n + 2 * g,) * 3, bool)
    m[g:g + n, g:g + n, g:g + n] = True
    s = slice(g, g + n)
    m[g - 1, s, s] = m[g + n, s, s] = True
    m[s, g - 1, s] = m[s, g + n, s] = True
    m[s, s, g - 1] = m[s, s, g + n] = True
    return m


def interior(a, n=N, g=G):
    return a[..., g:g + n, g:g + n, g:g + n]


class OracleGrid:
    """The oracle's view of a tree's block table."""

    def __init__(self, tree, g=G):
        self.tb = tree.table()
        self.tb["g"] = g
        self.tab = {k: v for k, v in self.tb.items()}
        self.tab["ftype"] = self.tb["ftype"].reshape(-1)
        self.tab["nbr"] = self.tb["nbr"].reshape(-1)
        self.tab["sub"] = self.tb["sub"].reshape(-1)
        self.tab["org"] = self.tb["org"].reshape(-1)
        self.grid = O.make_grid(self.tab)
        self.L = O.lib()
        self.nb = self.tb["nblk"]
        self.g = g
        self.shape = (self.nb, N + 2 * g, N + 2 * g, N + 2 * g)

    def zeros(self):
        return np.zeros(self.shape)

    def ghost_fill(self, f, updater, bct, bcv):
        self.L.orc_ghost_fill(ctypes.byref(self.grid), O.ptr(f), updater, O.cint6(bct), O.cdbl6(bcv))

    def ref(self):
        return ctypes.byref(self.grid)


def bc_lists(params, group):
    return [params.bct[group][f] for f in range(6)], [params.bcv[group][f] for f in range(6)]


def poisson_system(og, rng, cut=None):
    """A random diagonally dominant 7-band system on a grid (A bands as padded fields)."""
    A = [og.zeros() for _ in range(7)]
    for q in range(1, 7):
        A[q][:] = rng.uniform(0.5, 1.5, og.shape)
    A[0][:] = -(sum(A[q] for q in range(1, 7))) - rng.uniform(0.0, 0.5, og.shape)
    b = rng.uniform(-1, 1, og.shape)
    return A, b
