"""ctypes binding of libffvhc.so, the MI355X-native FFV-HC hot path (C ABI in include/ffvhc.h).

There is no CPU fallback: every call goes to the HIP library, and importing this module fails
if the library has not been built (`python -c "import __graft_entry__ as g; g.build()"`).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libffvhc.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libffvhc.so is not built (run __graft_entry__.build())")

_L = ctypes.CDLL(LIB_PATH)
_P = ctypes.POINTER
c_int, c_double, c_ll = ctypes.c_int, ctypes.c_double, ctypes.c_longlong

MAX_REGIONS = 32
X_M, X_P, Y_M, Y_P, Z_M, Z_P = range(6)
BC_UX, BC_UY, BC_UZ, BC_P, BC_NULL, BC_ZERO = range(6)
DIRICHLET, NEUMANN, PERIODIC, NONE = 0, 1, 9, -1


class Region(ctypes.Structure):
    _fields_ = [("kind", c_int), ("a", c_double * 3), ("b", c_double * 3), ("r", c_double), ("dr", c_double),
                ("level", c_int)]


class TreeSpec(ctypes.Structure):
    _fields_ = [("root", c_int * 3), ("origin", c_double * 3), ("root_length", c_double), ("periodic", c_int * 3),
                ("n", c_int), ("min_level", c_int), ("margin", c_double), ("ordering", c_int),
                ("nregions", c_int), ("regions", Region * MAX_REGIONS)]


class SphereSpec(ctypes.Structure):
    _fields_ = [("center", c_double * 3), ("radius", c_double), ("cut_id", c_int), ("cutoff", c_double),
                ("hole_filling", c_int * 3), ("seed", c_double * 3)]


class CaseParams(ctypes.Structure):
    _fields_ = [("rho", c_double), ("mu", c_double), ("dt", c_double), ("gx", c_double), ("gy", c_double),
                ("gz", c_double), ("betag", c_double), ("tr", c_double), ("bct", (c_int * 6) * 6),
                ("bcv", (c_double * 6) * 6)]


class StepFields(ctypes.Structure):
    """Field handles of one time step (ffvhc_step_fields); uc/ucp and ud/udp swap on return."""
    _fields_ = [("ux", c_int), ("uy", c_int), ("uz", c_int), ("p", c_int), ("uc", c_int * 3), ("ucp", c_int * 3),
                ("ud", c_int * 3), ("udp", c_int * 3), ("v", c_int * 6), ("A", c_int * 7), ("b", c_int),
                ("lapp", c_int), ("t", c_int), ("f", c_int * 3)]


class StepParams(ctypes.Structure):
    _fields_ = [("solver", c_int), ("mixed", c_int), ("omega", c_double), ("countMax", c_int), ("eps", c_double),
                ("countPre", c_int), ("inner_eps", c_double), ("ref_active", c_int), ("ref", c_double * 3),
                ("ref_p", c_double), ("u_implicit", c_int), ("alpha_u", c_double), ("omega_u", c_double),
                ("countMax_u", c_int), ("eps_u", c_double), ("mixed_method", c_int)]


class StepInfo(ctypes.Structure):
    _fields_ = [("countP", c_int), ("residualP", c_double), ("outer", c_int), ("status", c_int),
                ("countU", c_int * 3), ("residualU", c_double * 3)]


_vp = ctypes.c_void_p
_sig = {
    "ffvhc_last_error": (ctypes.c_char_p, []),
    "ffvhc_tree_build": (c_int, [_P(TreeSpec), _P(_vp)]),
    "ffvhc_tree_num_blocks": (c_int, [_vp]),
    "ffvhc_tree_max_level": (c_int, [_vp]),
    "ffvhc_tree_table": (c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffvhc_tree_free": (None, [_vp]),
    "ffvhc_geometry_sphere": (c_int, [_vp, _P(SphereSpec), c_int, _vp, _vp, _vp, _vp]),
    "ffvhc_ctx_create": (c_int, [c_int, c_ll, _P(_vp)]),
    "ffvhc_ctx_destroy": (None, [_vp]),
    "ffvhc_ctx_sync": (c_int, [_vp]),
    "ffvhc_ctx_device_bytes": (c_ll, [_vp]),
    "ffvhc_ctx_set_option": (c_int, [_vp, c_int, c_int]),
    "ffvhc_ctx_select": (c_int, [_vp, _vp, c_int]),
    "ffvhc_case_add": (c_int, [_vp, _vp, _P(CaseParams), c_int, _vp, _vp, _vp, _P(c_int)]),
    "ffvhc_case_remove": (c_int, [_vp, c_int]),
    "ffvhc_case_num_blocks": (c_int, [_vp, c_int]),
    "ffvhc_case_set_bc": (c_int, [_vp, c_int, c_int, c_int, c_int, c_double]),
    "ffvhc_field_alloc": (c_int, [_vp, c_int, _P(c_int)]),
    "ffvhc_field_free": (c_int, [_vp, c_int]),
    "ffvhc_field_upload": (c_int, [_vp, c_int, c_int, c_int, _vp]),
    "ffvhc_field_download": (c_int, [_vp, c_int, c_int, c_int, _vp]),
    "ffvhc_field_fill": (c_int, [_vp, c_int, c_double]),
    "ffvhc_field_copy": (c_int, [_vp, c_int, c_int]),
    "ffvhc_ghost_fill": (c_int, [_vp, c_int, c_int, c_int]),
    "ffvhc_rbgs": (c_int, [_vp, c_int, _vp, c_int, c_int, c_double, c_int, c_double, _vp, _vp]),
    "ffvhc_rbgs_fixed": (c_int, [_vp, c_int, _vp, c_int, c_int, c_double, c_int, _vp]),
    "ffvhc_rbgs_fixed_average": (c_int, [_vp, c_int, _vp, c_int, c_int, c_double, c_int, c_double, _vp, _vp]),
    "ffvhc_cg": (c_int, [_vp, c_int, _vp, c_int, c_int, c_double, c_int, c_int, c_double, _vp, _vp]),
    "ffvhc_bicgstab": (c_int, [_vp, c_int, _vp, c_int, c_int, c_double, c_int, c_int, c_double, _vp, _vp]),
    "ffvhc_bicgstab_timing": (c_int, [_vp, c_int, _vp, c_int, c_double, c_int, _vp]),
    "ffvhc_bicgstab_timing_mixed": (c_int, [_vp, c_int, c_int, _vp, c_int, c_double, c_int, _vp]),
    "ffvhc_coarse_operator": (c_int, [_vp, _vp, c_int, _vp, _vp]),
    "ffvhc_stencil_classes": (c_int, [_vp, _vp, _vp, _vp]),
    "ffvhc_cg_timing": (c_int, [_vp, c_int, _vp, c_int, c_double, c_int, _vp]),
    "ffvhc_jacobi": (c_int, [_vp, c_int, _vp, c_int, c_int, c_double, c_int, c_double, _vp, _vp]),
    "ffvhc_step": (c_int, [_vp, _P(StepFields), _P(StepParams), _vp, _vp]),
    "ffvhc_case_set_step": (c_int, [_vp, c_int, c_int, c_int]),
    "ffvhc_case_fill_phase": (c_int, [_vp, c_int, _vp, c_int, _vp]),
    "ffvhc_case_download_phase": (c_int, [_vp, c_int, c_int, _vp]),
    "ffvhc_force_p": (c_int, [_vp, c_int, c_int, _vp]),
    "ffvhc_force_p_noalloc": (c_int, [_vp, c_int, c_int, _vp]),
    "ffvhc_field_stats": (c_int, [_vp, c_int, _vp]),
    "ffvhc_cfl_check": (c_int, [_vp, c_int, _vp]),
    "ffvhc_field_dump": (c_int, [_vp, c_int, c_int, c_int, c_int, c_int, ctypes.c_char_p]),
    "ffvhc_field_load": (c_int, [_vp, c_int, c_int, ctypes.c_char_p]),
    "ffvhc_dump_header": (c_int, [ctypes.c_char_p, _vp]),
    "ffvhc_case_step": (c_int, [_vp, c_int]),
    "ffvhc_residual": (c_int, [_vp, c_int, _vp, c_int, _vp]),
    "ffvhc_solve_mixed": (c_int, [_vp, c_int, c_int, _vp, c_int, c_int, c_double, c_int, c_double, c_double, _vp,
                                  _vp, _vp]),
}
for _name, (_res, _args) in _sig.items():
    if hasattr(_L, _name):
        _f = getattr(_L, _name)
        _f.restype = _res
        _f.argtypes = _args


class FFVError(RuntimeError):
    def __init__(self, code, where):
        msg = _L.ffvhc_last_error()
        super().__init__("%s failed (%d): %s" % (where, code, msg.decode() if msg else ""))
        self.code = code


def check(rc, where):
    if rc != 0:
        raise FFVError(rc, where)
    return rc


def dump_header(path):
    """{nx, ny, nz, vc, nE, nB} of a Dump2 file."""
    h = np.zeros(6, np.int32)
    check(_L.ffvhc_dump_header(str(path).encode(), _ptr(h)), "ffvhc_dump_header")
    return h


def lib():
    return _L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------------------------ octree

def tree_spec(root=(1, 1, 1), origin=(0.0, 0.0, 0.0), root_length=1.0, periodic=(0, 0, 0), n=16, min_level=0,
              margin=0.0, ordering=1, regions=()):
    s = TreeSpec()
    for d in range(3):
        s.root[d] = int(root[d])
        s.origin[d] = float(origin[d])
        s.periodic[d] = int(periodic[d])
    s.root_length = float(root_length)
    s.n = int(n)
    s.min_level = int(min_level)
    s.margin = float(margin)
    s.ordering = int(ordering)
    s.nregions = len(regions)
    for i, r in enumerate(regions):
        g = s.regions[i]
        g.kind = {"ball": 0, "shell": 1, "box": 2}[r["kind"]]
        for d in range(3):
            g.a[d] = float(r.get("a", (0, 0, 0))[d])
            g.b[d] = float(r.get("b", (0, 0, 0))[d])
        g.r = float(r.get("r", 0.0))
        g.dr = float(r.get("dr", 0.0))
        g.level = int(r["level"])
    return s


class Tree:
    """Block octree built by the library (host code, no GPU needed)."""

    def __init__(self, **kw):
        self.spec = tree_spec(**kw)
        h = ctypes.c_void_p()
        check(_L.ffvhc_tree_build(ctypes.byref(self.spec), ctypes.byref(h)), "ffvhc_tree_build")
        self.h = h
        self.nblk = _L.ffvhc_tree_num_blocks(h)
        self.max_level = _L.ffvhc_tree_max_level(h)
        self.n = self.spec.n

    def table(self):
        nb = self.nblk
        t = {"nblk": nb, "n": self.n, "level": np.zeros(nb, np.int32), "org": np.zeros((nb, 3)),
             "dx": np.zeros(nb), "ftype": np.zeros((nb, 6), np.int32), "nbr": np.zeros((nb, 6, 4), np.int32),
             "sub": np.zeros((nb, 6), np.int32)}
        check(_L.ffvhc_tree_table(self.h, _ptr(t["level"]), _ptr(t["org"]), _ptr(t["dx"]), _ptr(t["ftype"]),
                                  _ptr(t["nbr"]), _ptr(t["sub"])), "ffvhc_tree_table")
        return t

    def geometry_sphere(self, center, radius, cut_id=1, cutoff=0.01, hole_filling=(1, 1, 0), seed=None, g=2):
        nb, n = self.nblk, self.n
        npad = (n + 2 * g) ** 3
        s = SphereSpec()
        for d in range(3):
            s.center[d] = float(center[d])
        s.radius = float(radius)
        s.cut_id = int(cut_id)
        s.cutoff = float(cutoff)
        for d in range(3):
            s.hole_filling[d] = int(hole_filling[d])
        if seed is None:
            seed = (self.spec.origin[0] + 1e-9, self.spec.origin[1] + 1e-9, self.spec.origin[2] + 1e-9)
        for d in range(3):
            s.seed[d] = float(seed[d])
        c = np.ones((6, nb, npad))
        cid = np.zeros((6, nb, npad), np.int32)
        pid = np.zeros((nb, npad), np.int32)
        counts = np.zeros(4, np.int64)
        check(_L.ffvhc_geometry_sphere(self.h, ctypes.byref(s), g, _ptr(c), _ptr(cid), _ptr(pid), _ptr(counts)),
              "ffvhc_geometry_sphere")
        return c, cid, pid, counts

    def __del__(self):
        if getattr(self, "h", None):
            _L.ffvhc_tree_free(self.h)
            self.h = None


# ----------------------------------------------------------------------------------- context

def case_params(rho=1.0, mu=0.01, dt=0.01, gravity=(0.0, 0.0, 0.0), betag=0.0, tr=0.0, bc=None):
    """bc: {group: [(type, value) x 6]}; default every group Neumann 0."""
    p = CaseParams()
    p.rho, p.mu, p.dt = rho, mu, dt
    p.gx, p.gy, p.gz = gravity
    p.betag, p.tr = betag, tr
    for gidx in range(6):
        for f in range(6):
            p.bct[gidx][f] = NEUMANN if gidx != BC_ZERO else DIRICHLET
            p.bcv[gidx][f] = 0.0
    for gidx, faces in (bc or {}).items():
        for f, (t, v) in enumerate(faces):
            p.bct[gidx][f] = int(t)
            p.bcv[gidx][f] = float(v)
    return p


class Context:
    def __init__(self, max_blocks, device=0):
        h = ctypes.c_void_p()
        check(_L.ffvhc_ctx_create(int(device), int(max_blocks), ctypes.byref(h)), "ffvhc_ctx_create")
        self.h = h
        self.cases = {}

    def close(self):
        if self.h:
            _L.ffvhc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sync(self):
        check(_L.ffvhc_ctx_sync(self.h), "ffvhc_ctx_sync")

    def device_bytes(self):
        return _L.ffvhc_ctx_device_bytes(self.h)

    def add_case(self, tree, params, geometry=None, g=2):
        cid = c_int()
        c, ci, pi = (None, None, None) if geometry is None else geometry[:3]
        check(_L.ffvhc_case_add(self.h, tree.h, ctypes.byref(params), g, _ptr(c), _ptr(ci), _ptr(pi),
                                ctypes.byref(cid)), "ffvhc_case_add")
        self.cases[cid.value] = tree
        return cid.value

    def remove_case(self, case_id):
        check(_L.ffvhc_case_remove(self.h, int(case_id)), "ffvhc_case_remove")
        self.cases.pop(case_id, None)

    def fill_phase(self, case_id, seed, cut_id=1):
        """Device flood fill of the case's phase from the cell holding `seed`; (fluid, solid) counts."""
        sd = np.asarray(seed, np.float64)
        counts = np.zeros(2, np.int64)
        check(_L.ffvhc_case_fill_phase(self.h, int(case_id), _ptr(sd), int(cut_id), _ptr(counts)),
              "ffvhc_case_fill_phase")
        return counts

    def download_phase(self, case_id, g=2):
        """The case's phase (1 fluid, 0 solid) as padded int blocks."""
        nb = self.cases[case_id].nblk
        out = np.zeros((nb, 16 + 2 * g, 16 + 2 * g, 16 + 2 * g), np.int32)
        check(_L.ffvhc_case_download_phase(self.h, int(case_id), int(g), _ptr(out)), "ffvhc_case_download_phase")
        return out

    def set_graphs(self, on):
        """Replay the BiCGSTAB iteration as a captured HIP graph, or launch one by one (default)."""
        check(_L.ffvhc_ctx_set_option(self.h, 1, int(bool(on))), "ffvhc_ctx_set_option")

    def set_shadow_hash(self, on):
        """Seeded hash shadow residual for the compact-record solve methods (11-28), or r0 = r
        (default)."""
        check(_L.ffvhc_ctx_set_option(self.h, 3, int(bool(on))), "ffvhc_ctx_set_option")

    def set_setup_reuse(self, on):
        """Let ffvhc_step keep its mixed-precision pressure setup from step to step while the
        matrix cannot have changed (default), or make it every step."""
        check(_L.ffvhc_ctx_set_option(self.h, 4, int(bool(on))), "ffvhc_ctx_set_option")

    def set_coarse_gemm_min(self, nblk):
        """Smallest case (blocks) whose coarse operator is inverted with GEMM panels (default 2048)."""
        check(_L.ffvhc_ctx_set_option(self.h, 5, int(nblk)), "ffvhc_ctx_set_option")

    def set_fuse_fin(self, on):
        """Let the multiplicative two-level methods (23-28) form each iteration's end (convergence
        test, next rho, beta) in the next iteration's first coarse application (default), or give
        it a launch of its own every iteration."""
        check(_L.ffvhc_ctx_set_option(self.h, 6, int(bool(on))), "ffvhc_ctx_set_option")

    def set_classes(self, on):
        """Let the compact-record solve methods (11-28) read the bands of blocks that are uniform
        per stencil position class from 27 rows (default), or read every block per cell."""
        check(_L.ffvhc_ctx_set_option(self.h, 2, int(bool(on))), "ffvhc_ctx_set_option")

    def select(self, ids=None):
        if ids is None:
            check(_L.ffvhc_ctx_select(self.h, None, 0), "ffvhc_ctx_select")
        else:
            a = np.asarray(ids, np.int32)
            check(_L.ffvhc_ctx_select(self.h, _ptr(a), len(a)), "ffvhc_ctx_select")

    def set_bc(self, case_id, group, face, type_, value):
        check(_L.ffvhc_case_set_bc(self.h, case_id, group, face, type_, value), "ffvhc_case_set_bc")

    def field(self, prec=8):
        f = c_int()
        check(_L.ffvhc_field_alloc(self.h, prec, ctypes.byref(f)), "ffvhc_field_alloc")
        return f.value

    def free(self, f):
        check(_L.ffvhc_field_free(self.h, f), "ffvhc_field_free")

    def upload(self, f, case_id, padded, g=2):
        a = np.ascontiguousarray(padded)
        check(_L.ffvhc_field_upload(self.h, f, case_id, g, _ptr(a)), "ffvhc_field_upload")

    def download(self, f, case_id, g=2, out=None, dtype=np.float64):
        nb = _L.ffvhc_case_num_blocks(self.h, case_id)
        n = 16
        if out is None:
            out = np.zeros((nb, n + 2 * g, n + 2 * g, n + 2 * g), dtype)
        check(_L.ffvhc_field_download(self.h, f, case_id, g, _ptr(out)), "ffvhc_field_download")
        return out

    def fill(self, f, v):
        check(_L.ffvhc_field_fill(self.h, f, float(v)), "ffvhc_field_fill")

    def copy(self, dst, src):
        check(_L.ffvhc_field_copy(self.h, dst, src), "ffvhc_field_copy")

    def ghost_fill(self, f, group, updater=1):
        check(_L.ffvhc_ghost_fill(self.h, f, group, updater), "ffvhc_ghost_fill")

    def _nsel(self):
        return len(self.cases)

    def rbgs(self, x, A, b, group, omega, count_max, eps, ncases=None):
        nc = ncases or self._nsel()
        cnt = np.zeros(nc, np.int32)
        res = np.zeros(nc)
        a = np.asarray(A, np.int32)
        rc = _L.ffvhc_rbgs(self.h, x, _ptr(a), b, group, omega, count_max, eps, _ptr(cnt), _ptr(res))
        if rc < 0:
            check(rc, "ffvhc_rbgs")
        return cnt, res, rc

    def rbgs_fixed(self, x, A, b, group, omega, nsweep, with_finalize=False):
        """nsweep fused RBGS iterations; device ms of the iteration kernels (and of the per-case
        residual finalisations when with_finalize)."""
        ms = np.zeros(2)
        a = np.asarray(A, np.int32)
        check(_L.ffvhc_rbgs_fixed(self.h, x, _ptr(a), b, group, omega, nsweep, _ptr(ms)), "ffvhc_rbgs_fixed")
        return (ms[0], ms[1]) if with_finalize else ms[0]

    def jacobi(self, x, A, b, group, omega, count_max, eps, ncases=None):
        nc = ncases or self._nsel()
        cnt = np.zeros(nc, np.int32)
        res = np.zeros(nc)
        a = np.asarray(A, np.int32)
        rc = _L.ffvhc_jacobi(self.h, x, _ptr(a), b, group, omega, count_max, eps, _ptr(cnt), _ptr(res))
        if rc < 0:
            check(rc, "ffvhc_jacobi")
        return cnt, res, rc

    def rbgs_fixed_average(self, x, A, b, group, omega, count_max, eps, ncases=None):
        nc = ncases or self._nsel()
        cnt = np.zeros(nc, np.int32)
        res = np.zeros(nc)
        a = np.asarray(A, np.int32)
        rc = _L.ffvhc_rbgs_fixed_average(self.h, x, _ptr(a), b, group, omega, count_max, eps, _ptr(cnt), _ptr(res))
        if rc < 0:
            check(rc, "ffvhc_rbgs_fixed_average")
        return cnt, res, rc

    def cg(self, x, A, b, group, omega, count_pre, count_max, eps, ncases=None):
        nc = ncases or self._nsel()
        cnt = np.zeros(nc, np.int32)
        res = np.zeros(nc)
        a = np.asarray(A, np.int32)
        rc = _L.ffvhc_cg(self.h, x, _ptr(a), b, group, omega, count_pre, count_max, eps, _ptr(cnt), _ptr(res))
        if rc < 0:
            check(rc, "ffvhc_cg")
        return cnt, res, rc

    def bicgstab(self, x, A, b, group, omega, count_pre, count_max, eps, ncases=None):
        nc = ncases or self._nsel()
        cnt = np.zeros(nc, np.int32)
        res = np.zeros(nc)
        a = np.asarray(A, np.int32)
        rc = _L.ffvhc_bicgstab(self.h, x, _ptr(a), b, group, omega, count_pre, count_max, eps, _ptr(cnt), _ptr(res))
        if rc < 0:
            check(rc, "ffvhc_bicgstab")
        return cnt, res, rc

    def bicgstab_timing(self, x, A, b, omega, niter, method=None):
        """Device ms of the five BiCGSTAB kernels, of the finalisations and of the coarse
        corrections, summed over niter (method 5-22: the kernels of that solve_mixed method)."""
        ms = np.zeros(7)
        a = np.asarray(A, np.int32)
        if method is None:
            check(_L.ffvhc_bicgstab_timing(self.h, x, _ptr(a), b, omega, niter, _ptr(ms)), "ffvhc_bicgstab_timing")
        else:
            check(_L.ffvhc_bicgstab_timing_mixed(self.h, method, x, _ptr(a), b, omega, niter, _ptr(ms)),
                  "ffvhc_bicgstab_timing_mixed")
        return ms

    def cg_timing(self, x, A, b, omega, niter):
        """Device ms of the three PCG kernels and of the finalisations, summed over niter."""
        ms = np.zeros(4)
        a = np.asarray(A, np.int32)
        check(_L.ffvhc_cg_timing(self.h, x, _ptr(a), b, omega, niter, _ptr(ms)), "ffvhc_cg_timing")
        return ms

    def step_fields(self, with_t=False, with_f=False):
        """Allocate every FP64 field of a time step; returns a StepFields struct."""
        F = StepFields()
        F.ux, F.uy, F.uz, F.p = (self.field(8) for _ in range(4))
        for d in range(3):
            F.uc[d], F.ucp[d], F.ud[d], F.udp[d] = (self.field(8) for _ in range(4))
        for q in range(6):
            F.v[q] = self.field(8)
        for q in range(7):
            F.A[q] = self.field(8)
        F.b, F.lapp = self.field(8), self.field(8)
        F.t = self.field(8) if with_t else -1
        for d in range(3):
            F.f[d] = self.field(8) if with_f else -1
        return F

    def step(self, F, solver=0, omega=1.2, count_max=100, eps=1e-8, count_pre=1, mixed=False, inner_eps=1e-5,
             ref=None, implicit_u=False, alpha_u=0.5, omega_u=1.0, count_max_u=100, eps_u=1e-8, ncases=None,
             mixed_method=-1):
        """One Solver::Update step of the selected cases; F's uc/ucp, ud/udp swap in place.
        mixed_method (with mixed): the solve_mixed method for P (-1: chosen by solver).
        Returns (per-case StepInfo list, stage ms [7], status)."""
        P = StepParams()
        P.solver, P.mixed, P.omega, P.countMax, P.eps = solver, int(mixed), omega, count_max, eps
        P.mixed_method = mixed_method
        P.countPre, P.inner_eps = count_pre, inner_eps
        P.u_implicit, P.alpha_u, P.omega_u, P.countMax_u, P.eps_u = int(implicit_u), alpha_u, omega_u, count_max_u, eps_u
        if ref is not None:
            P.ref_active = 1
            P.ref[0], P.ref[1], P.ref[2] = ref[:3]
            P.ref_p = ref[3]
        nc = ncases or self._nsel()
        info = (StepInfo * nc)()
        ms = np.zeros(7)
        rc = _L.ffvhc_step(self.h, ctypes.byref(F), ctypes.byref(P), info, _ptr(ms))
        if rc < 0:
            check(rc, "ffvhc_step")
        return list(info), ms, rc

    def force_p(self, p, cut_id, ncases=None):
        """Pressure force (fx, fy, fz) on the faces with boundary id cut_id, per selected case."""
        f = np.zeros((ncases or self._nsel(), 3))
        check(_L.ffvhc_force_p(self.h, p, cut_id, _ptr(f)), "ffvhc_force_p")
        return f

    def force_p_noalloc(self, p, cut_id, ncases=None):
        """force_p through the allocation-free entry point (fixed per-device scratch)."""
        f = np.zeros((ncases or self._nsel(), 3))
        check(_L.ffvhc_force_p_noalloc(self.h, p, cut_id, _ptr(f)), "ffvhc_force_p_noalloc")
        return f

    def coarse_operator(self, A, case_id):
        """(A_c, its device inverse) of the two-level preconditioner for one case."""
        n = self.cases[case_id].nblk
        ac, ainv = np.zeros((n, n)), np.zeros((n, n))
        a = np.asarray(A, np.int32)
        check(_L.ffvhc_coarse_operator(self.h, _ptr(a), case_id, _ptr(ac), _ptr(ainv)), "ffvhc_coarse_operator")
        return ac, ainv

    def stencil_classes(self, A):
        """(class-uniform blocks, other blocks) of the selected cases for the FP64 bands A."""
        a = np.asarray(A, np.int32)
        nu, no = c_int(), c_int()
        check(_L.ffvhc_stencil_classes(self.h, _ptr(a), ctypes.byref(nu), ctypes.byref(no)), "ffvhc_stencil_classes")
        return nu.value, no.value

    def force_v(self, ux, uy, uz, cut_id, ncases=None):
        """Viscous force (fx, fy, fz) on the faces with boundary id cut_id, per selected case."""
        f = np.zeros((ncases or self._nsel(), 3))
        check(_L.ffvhc_force_v(self.h, ux, uy, uz, cut_id, _ptr(f)), "ffvhc_force_v")
        return f

    def stats(self, f, ncases=None):
        """(sum, max, min, |max|, |min|) over the interior of each selected case."""
        out = np.zeros((ncases or self._nsel(), 5))
        check(_L.ffvhc_field_stats(self.h, f, _ptr(out)), "ffvhc_field_stats")
        return out

    def cfl(self, ux, ncases=None):
        """CFL numbers of the reference's guard per selected case and the status (2 = tripped)."""
        out = np.zeros(ncases or self._nsel())
        rc = _L.ffvhc_cfl_check(self.h, ux, _ptr(out))
        if rc < 0:
            check(rc, "ffvhc_cfl_check")
        return out, rc

    def dump(self, f, case_id, path, g=2, group=BC_P, updater=1):
        """Write a field of a case in the reference's Dump2 layout, ghosts filled first with the BC
        group and updater (1 for p, 5 for the velocity components)."""
        check(_L.ffvhc_field_dump(self.h, f, case_id, g, group, updater, str(path).encode()), "ffvhc_field_dump")

    def load(self, f, case_id, path):
        """Read a Dump2 file into a field of a case (ghosts: call ghost_fill afterwards)."""
        check(_L.ffvhc_field_load(self.h, f, case_id, str(path).encode()), "ffvhc_field_load")

    def set_step(self, case_id, step, accel_steps=0):
        check(_L.ffvhc_case_set_step(self.h, case_id, step, accel_steps), "ffvhc_case_set_step")

    def residual(self, x, A, b, ncases=None):
        """FP64 ||b - Ax|| / ||b|| per selected case (x's ghost layer must be current)."""
        res = np.zeros(ncases or self._nsel())
        a = np.asarray(A, np.int32)
        check(_L.ffvhc_residual(self.h, x, _ptr(a), b, _ptr(res)), "ffvhc_residual")
        return res

    def solve_mixed(self, method, x, A, b, group, omega, count_max, eps, inner_eps, ncases=None):
        nc = ncases or self._nsel()
        cnt = np.zeros(nc, np.int32)
        res = np.zeros(nc)
        outer = np.zeros(nc, np.int32)
        a = np.asarray(A, np.int32)
        rc = _L.ffvhc_solve_mixed(self.h, method, x, _ptr(a), b, group, omega, count_max, eps, inner_eps, _ptr(cnt),
                                  _ptr(res), _ptr(outer))
        if rc < 0:
            check(rc, "ffvhc_solve_mixed")
        return cnt, res, outer, rc
