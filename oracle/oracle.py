"""ctypes access to the CPU oracle and to the reference Fortran build (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The oracle (liboracle.so) restates the reference algorithm in C; oracle/_ref holds the
reference's own Fortran kernels compiled from /root/reference by oracle/Makefile.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.POINTER
dp = _P(ctypes.c_double)
ip = _P(ctypes.c_int)


def build(quiet=True):
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = ctypes.CDLL(path)
        _declare(_lib)
    return _lib


def ref_path(kind="d8"):
    """Path of the reference Fortran kernels of `kind`, or None when oracle/_ref is not built.
    Checking it maps nothing into the process."""
    name = {"d8": "libffvref.so", "q16": "libffvref_q.so"}[kind]
    path = os.path.join(HERE, "_ref", name)
    return path if os.path.exists(path) else None


def ref(kind="d8"):
    """Reference Fortran kernels: kind 'd8' (all-double literals) or 'q16' (as configured)."""
    path = ref_path(kind)
    return None if path is None else ctypes.CDLL(path)


def ptr(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(ip)
    raise TypeError(a.dtype)


def parr(arrs, ctype=ctypes.c_double):
    """C array of pointers (real* const* / int* const*)."""
    T = _P(ctype) * len(arrs)
    return T(*[a.ctypes.data_as(_P(ctype)) for a in arrs])


class Grid(ctypes.Structure):
    _fields_ = [("nblk", ctypes.c_int), ("n", ctypes.c_int), ("g", ctypes.c_int),
                ("ftype", ip), ("nbr", ip), ("sub", ip), ("dx", dp), ("org", dp)]


class Fields(ctypes.Structure):
    _fields_ = [("ux", dp), ("uy", dp), ("uz", dp), ("p", dp),
                ("uc", dp * 3), ("ucp", dp * 3), ("ud", dp * 3), ("udp", dp * 3),
                ("v", dp * 6), ("A", dp * 7), ("b", dp), ("lapp", dp),
                ("t0", dp), ("f", dp * 3), ("cgs", dp * 5), ("bis", dp * 9),
                ("c", dp * 6), ("cid", ip * 6), ("pid", ip)]


class Params(ctypes.Structure):
    _fields_ = [("rho", ctypes.c_double), ("mu", ctypes.c_double), ("dt", ctypes.c_double),
                ("gx", ctypes.c_double), ("gy", ctypes.c_double), ("gz", ctypes.c_double),
                ("betag", ctypes.c_double), ("tr", ctypes.c_double),
                ("bct_u", (ctypes.c_int * 6) * 3), ("bcv_u", (ctypes.c_double * 6) * 3),
                ("bct_p", ctypes.c_int * 6), ("bcv_p", ctypes.c_double * 6),
                ("solver", ctypes.c_int), ("omega", ctypes.c_double), ("countMax", ctypes.c_int),
                ("eps", ctypes.c_double), ("countPre", ctypes.c_int), ("ncells", ctypes.c_int),
                ("accel_steps", ctypes.c_int),
                ("ref_active", ctypes.c_int), ("ref_x", ctypes.c_double), ("ref_y", ctypes.c_double),
                ("ref_z", ctypes.c_double), ("ref_p", ctypes.c_double),
                ("u_implicit", ctypes.c_int), ("alpha_u", ctypes.c_double), ("omega_u", ctypes.c_double),
                ("eps_u", ctypes.c_double), ("countMax_u", ctypes.c_int)]


class StepInfo(ctypes.Structure):
    _fields_ = [("countP", ctypes.c_int), ("residualP", ctypes.c_double), ("status", ctypes.c_int),
                ("times", ctypes.c_double * 7), ("countU", ctypes.c_int * 3), ("residualU", ctypes.c_double * 3)]


def _declare(L):
    d, i = ctypes.c_double, ctypes.c_int
    L.orc_calc_r2.restype = d
    L.orc_dot.restype = d
    L.orc_sf3d_calc_sum.restype = d
    L.orc_grid_dot.restype = d
    L.orc_grid_r2.restype = d
    L.orc_grid_sum.restype = d
    L.orc_bstl_detect_zerocut.restype = i
    L.orc_bstl_fill_holes_v1.restype = i
    L.orc_bstl_fill_holes_v2.restype = i
    for name in ["orc_rbgs", "orc_rbgs_fixed_average", "orc_jacobi", "orc_cg", "orc_bicgstab"]:
        getattr(L, name).restype = i


def make_grid(table):
    """Grid structure from a block-table dict (numpy arrays kept alive on the dict)."""
    g = Grid()
    g.nblk = int(table["nblk"])
    g.n = int(table["n"])
    g.g = int(table["g"])
    for key in ("ftype", "nbr", "sub"):
        table["_" + key] = np.ascontiguousarray(table[key], dtype=np.int32)
    for key in ("dx", "org"):
        table["_" + key] = np.ascontiguousarray(table[key], dtype=np.float64)
    g.ftype = ptr(table["_ftype"])
    g.nbr = ptr(table["_nbr"])
    g.sub = ptr(table["_sub"])
    g.dx = ptr(table["_dx"])
    g.org = ptr(table["_org"])
    table["_grid"] = g
    return g


def cint6(vals):
    return (ctypes.c_int * 6)(*[int(v) for v in vals])


def cdbl6(vals):
    return (ctypes.c_double * 6)(*[float(v) for v in vals])
