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


def geometry(og, rng):
    c = np.ones((6,) + og.shape)
    cid = np.zeros((6,) + og.shape, np.int32)
    for f in range(6):
        m = rng.random(og.shape) < 0.04
        cid[f][m] = rng.integers(1, 3, int(m.sum()))
        c[f][m] = rng.uniform(0.05, 1.0, int(m.sum())).astype(np.float32).astype(np.float64)
    pid = np.ones(og.shape, np.int32)
    pid[rng.random(og.shape) < 0.03] = 0
    return c, cid, pid


class OracleStep:
    """orc_update on host arrays mirroring the device fields."""

    def __init__(self, og, params, geo, arrays, solver, omega, count_max, eps, implicit=None):
        self.og = og
        self.a = arrays
        self.geo = geo
        F = O.Fields()
        for name in ("ux", "uy", "uz", "p", "b", "lapp"):
            setattr(F, name, O.ptr(arrays[name]))
        for d in range(3):
            F.uc[d], F.ucp[d] = O.ptr(arrays["uc"][d]), O.ptr(arrays["ucp"][d])
            F.ud[d], F.udp[d] = O.ptr(arrays["ud"][d]), O.ptr(arrays["udp"][d])
            F.f[d] = O.ptr(arrays["f"][d])
        for q in range(6):
            F.v[q] = O.ptr(arrays["v"][q])
            F.c[q] = O.ptr(geo[0][q])
            F.cid[q] = O.ptr(geo[1][q])
        for q in range(7):
            F.A[q] = O.ptr(arrays["A"][q])
        for q in range(5):
            F.cgs[q] = O.ptr(arrays["cgs"][q])
        for q in range(9):
            F.bis[q] = O.ptr(arrays["bis"][q])
        F.t0 = O.ptr(arrays["t"])
        F.pid = O.ptr(geo[2])
        self.F = F
        P = O.Params()
        P.rho, P.mu, P.dt = params.rho, params.mu, params.dt
        P.gx, P.gy, P.gz, P.betag, P.tr = params.gx, params.gy, params.gz, params.betag, params.tr
        for d in range(3):
            for f in range(6):
                P.bct_u[d][f] = params.bct[d][f]
                P.bcv_u[d][f] = params.bcv[d][f]
        for f in range(6):
            P.bct_p[f] = params.bct[ffvhc.BC_P][f]
            P.bcv_p[f] = params.bcv[ffvhc.BC_P][f]
        P.solver = {0: 0, 1: 1, 2: 2, 3: 3}[solver]
        P.omega, P.countMax, P.eps, P.countPre = omega, count_max, eps, 1
        P.ncells = og.nb * N ** 3
        if implicit is not None:
            P.u_implicit = 1
            P.alpha_u, P.omega_u, P.eps_u, P.countMax_u = implicit
        self.P = P

    def step(self, n):
        info = O.StepInfo()
        self.og.L.orc_update(self.og.ref(), ctypes.byref(self.F), ctypes.byref(self.P), n, ctypes.byref(info))
        return info


def initial_arrays(og, rng):
    z = og.zeros
    a = {"ux": 1.0 + 0.1 * rng.standard_normal(og.shape), "uy": 0.1 * rng.standard_normal(og.shape),
         "uz": 0.1 * rng.standard_normal(og.shape), "p": 0.05 * rng.standard_normal(og.shape),
         "b": z(), "lapp": z(), "t": z(),
         "uc": [z() for _ in range(3)], "ucp": [z() for _ in range(3)],
         "ud": [z() for _ in range(3)], "udp": [z() for _ in range(3)],
         "f": [z() for _ in range(3)], "v": [z() for _ in range(6)], "A": [z() for _ in range(7)],
         "cgs": [z() for _ in range(5)], "bis": [z() for _ in range(9)]}
    # face velocities consistent with the initial field (what calc_ab_p of step -1 would leave)
    a["v"][0][:] = a["ux"]
    a["v"][1][:] = a["ux"]
    return a
