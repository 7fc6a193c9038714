"""Per-block kernel cases shared by the oracle-vs-reference tests and the golden-fixture script.

Each case builds seeded inputs on one padded block, runs either the oracle (C restatement)
or the reference Fortran (oracle/_ref, CGMfuncs calling convention: every argument by
reference, trailing underscore) and returns the named outputs.

Inputs come from a counter-based generator (splitmix64) written out here, so the same
arrays are produced on any machine and numpy version.
"""
import ctypes

import numpy as np

from oracle import oracle as O

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix(seed, n):
    """n uniform doubles in [0,1) from splitmix64(seed*2^32 + i)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(32)) + np.arange(n, dtype=np.uint64)
        z = z * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Gen:
    def __init__(self, seed, sz, g):
        self.seed = seed * 1000
        self.sz = tuple(sz)
        self.g = g
        self.shape = tuple(s + 2 * g for s in sz)[::-1]  # (k, j, i), x fastest

    def _u(self):
        self.seed += 1
        return splitmix(self.seed, int(np.prod(self.shape))).reshape(self.shape)

    def real(self, lo=-1.0, hi=1.0):
        return np.ascontiguousarray(lo + (hi - lo) * self._u())

    def int(self, vals, probs):
        u = self._u()
        out = np.zeros(self.shape, dtype=np.int32)
        acc = 0.0
        for v, p in zip(vals, probs):
            out[(u >= acc) & (u < acc + p)] = v
            acc += p
        return np.ascontiguousarray(out)

    def cuts(self, cut):
        """c0..c5 as float32-representable distances in (0,1], cid0..5, pid."""
        c, cid = [], []
        for _ in range(6):
            d = self._u().astype(np.float32).astype(np.float64)
            d = np.where(d <= 0.0, 0.5, d)
            if cut:
                ids = self.int([0, 1, 2], [0.55, 0.30, 0.15])
            else:
                ids = np.zeros(self.shape, np.int32)
            c.append(np.ascontiguousarray(np.where(ids != 0, d, 1.0)))
            cid.append(ids)
        pid = self.int([1, 0], [0.9, 0.1]) if cut else np.ones(self.shape, np.int32)
        return c, cid, pid


D = ctypes.c_double
I = ctypes.c_int


def _r(x):
    return ctypes.byref(D(x))


def _i(x):
    return ctypes.byref(I(x))


def _sz(sz):
    return (I * 3)(*sz)


FACE_X = {0: "x3", 1: "x1", 2: "x4", 3: "x2", 4: "x6", 5: "x5"}  # BlockScalar3D.h:201-254
FACE_A = {0: "aw", 1: "ae", 2: "as", 3: "an", 4: "ab", 5: "at"}


def run_case(name, seed, sz, g, impl, cut=True):
    """Run one kernel case.  impl: 'oracle' or a reference CDLL.  Returns {output: array}."""
    G = Gen(seed, sz, g)
    L = O.lib()
    F = None if impl == "oracle" else impl
    s3 = _sz(sz)
    P = O.ptr
    A = [G.real(-1.0, 1.0) for _ in range(7)]
    A[0] = G.real(-7.0, -5.0)
    b = G.real()
    x = G.real()
    y = G.real()
    c, cid, pid = G.cuts(cut)
    out = {}
    if name == "jacobi_smoother":
        x1 = G.real()
        if F is None:
            L.orc_jacobi_smoother(P(x1), P(x), *[P(a) for a in A], P(b), D(1.1), s3, I(g))
        else:
            F.jacobi_smoother_(P(x1), P(x), *[P(a) for a in A], P(b), _r(1.1), s3, _i(g))
        out["x1"] = x1
    elif name == "rbgs_smoother":
        for color in (0, 1):
            if F is None:
                L.orc_rbgs_smoother(P(x), *[P(a) for a in A], P(b), D(1.2), I(color), I(0), s3, I(g))
            else:
                F.rbgs_smoother_(P(x), *[P(a) for a in A], P(b), _r(1.2), _i(color), _i(0), s3, _i(g))
        out["x"] = x
    elif name == "calc_ax":
        ax = G.real()
        if F is None:
            L.orc_calc_ax(P(ax), *[P(a) for a in A], P(x), s3, I(g))
        else:
            F.calc_ax_(P(ax), *[P(a) for a in A], P(x), s3, _i(g))
        out["ax"] = ax
    elif name == "calc_r":
        r = G.real()
        if F is None:
            L.orc_calc_r(P(r), *[P(a) for a in A], P(x), P(b), s3, I(g))
        else:
            F.calc_r_(P(r), *[P(a) for a in A], P(x), P(b), s3, _i(g))
        out["r"] = r
    elif name == "calc_r2":
        if F is None:
            v = L.orc_calc_r2(*[P(a) for a in A], P(x), P(b), s3, I(g))
        else:
            rr = D(0.0)
            F.calc_r2_(ctypes.byref(rr), *[P(a) for a in A], P(x), P(b), s3, _i(g))
            v = rr.value
        out["rr"] = np.array([v])
    elif name == "blas":
        ys = [y.copy() for _ in range(6)]
        s = {}
        if F is None:
            L.orc_fill(P(ys[0]), D(0.75), s3, I(g))
            L.orc_copy(P(ys[1]), P(x), s3, I(g))
            L.orc_adda(P(ys[2]), D(-0.3), s3, I(g))
            L.orc_axpy(P(ys[3]), P(x), D(0.7), s3, I(g))
            L.orc_xpay(P(ys[4]), P(x), D(-1.3), s3, I(g))
            z = b.copy()
            L.orc_axpyz(P(z), P(x), P(y), D(0.45), s3, I(g))
            zz = A[1].copy()
            L.orc_axpbypz(P(zz), P(x), P(y), D(0.45), D(-2.5), s3, I(g))
            s["dot"] = L.orc_dot(P(x), P(y), s3, I(g))
            s["sum"] = L.orc_sf3d_calc_sum(P(x), s3, I(g))
        else:
            F.fill_(P(ys[0]), _r(0.75), s3, _i(g))
            F.copy_(P(ys[1]), P(x), s3, _i(g))
            F.adda_(P(ys[2]), _r(-0.3), s3, _i(g))
            F.axpy_(P(ys[3]), P(x), _r(0.7), s3, _i(g))
            F.xpay_(P(ys[4]), P(x), _r(-1.3), s3, _i(g))
            z = b.copy()
            F.axpyz_(P(z), P(x), P(y), _r(0.45), s3, _i(g))
            zz = A[1].copy()
            F.axpbypz_(P(zz), P(x), P(y), _r(0.45), _r(-2.5), s3, _i(g))
            dv, sv = D(0.0), D(0.0)
            F.dot_(ctypes.byref(dv), P(x), P(y), s3, _i(g))
            F.sf3d_calc_sum_(ctypes.byref(sv), P(x), s3, _i(g))
            s["dot"], s["sum"] = dv.value, sv.value
        for q in range(5):
            out["y%d" % q] = ys[q]
        out["axpyz"] = z
        out["axpbypz"] = zz
        out["dot"] = np.array([s["dot"]])
        out["sum"] = np.array([s["sum"]])
    elif name.startswith("sup_"):
        # random valid index setups (coarse-fine geometry of Scalar3DUpdater1)
        kind = name[4:]
        rng = np.random.default_rng(seed)
        dst = G.real()
        nx = sz[0]
        for trial in range(4):
            if kind == "copy":
                sc = [int(rng.integers(1, s + 1)) for s in sz]
                i1s = [int(rng.integers(1, s - c + 2)) for s, c in zip(sz, sc)]
                i1d = [int(rng.integers(1 - g, s + g - c + 2)) for s, c in zip(sz, sc)]
            elif kind == "c2f":
                sc = [int(rng.integers(1, s + 1)) for s in sz]
                i1s = [int(rng.integers(1, 2 * s - c + 2)) for s, c in zip(sz, sc)]
                i1d = [int(rng.integers(1 - g, s + g - c + 2)) for s, c in zip(sz, sc)]
            else:
                sc = [int(rng.integers(1, s // 2 + 1)) for s in sz]
                i1s = [int(rng.integers(1, s // 2 - c + 2)) for s, c in zip(sz, sc)]
                i1d = [int(rng.integers(1 - g, s + g - c + 2)) for s, c in zip(sz, sc)]
            a_d, a_s, a_c = (I * 3)(*i1d), (I * 3)(*i1s), (I * 3)(*sc)
            fn = {"copy": "sup_copy_from_neighbor", "c2f": "sup_copy_from_neighbor_c2f",
                  "f2c": "sup_copy_from_neighbor_f2c"}[kind]
            if F is None:
                getattr(L, "orc_" + fn)(P(dst), a_d, P(x), a_s, a_c, s3, I(g))
            else:
                getattr(F, fn + "_")(P(dst), a_d, P(x), a_s, a_c, s3, _i(g))
        out["dst"] = dst
        del nx
    elif name == "bc_x":
        for face in range(6):
            for typ in (0, 1, 9):
                if typ == 9 and face % 2 == 0:
                    continue
                f = G.real()
                xc = 0.37 + 0.1 * face
                if F is None:
                    L.orc_bc_x(P(f), I(face), I(typ), D(xc), s3, I(g))
                else:
                    sfx = {0: "d", 1: "n", 9: "p"}[typ]
                    getattr(F, "bc_%s_%s_" % (FACE_X[face], sfx))(P(f), _r(xc), s3, _i(g))
                out["f%d_%d" % (face, typ)] = f
    elif name == "bc_a":
        for face in range(6):
            for typ in (0, 1):
                Ap, An_, bb = A[0].copy(), A[1 + face].copy(), b.copy()
                xc = -0.21 + 0.05 * face
                if F is None:
                    L.orc_bc_a(P(Ap), P(An_), P(bb), I(face), I(typ), D(xc), s3, I(g))
                else:
                    sfx = {0: "d", 1: "n"}[typ]
                    getattr(F, "bc_%s_%s_" % (FACE_A[face], sfx))(P(Ap), P(An_), P(bb), _r(xc), s3, _i(g))
                out["Ap%d_%d" % (face, typ)] = Ap
                out["An%d_%d" % (face, typ)] = An_
                out["b%d_%d" % (face, typ)] = bb
    else:
        out.update(_run_bcut(name, G, L, F, s3, g, A, b, x, y, c, cid, pid))
    return out


def _run_bcut(name, G, L, F, s3, g, A, b, x, y, c, cid, pid):
    P = O.ptr
    out = {}
    cp = [P(a) for a in c]
    ip_ = [P(a) for a in cid]
    cpp = O.parr(c)
    ipp = O.parr(cid, ctypes.c_int)
    v = [G.real() for _ in range(6)]
    dx, dt, rho, mu = 0.125, 0.01, 1.3, 0.02
    if name == "calc_c_f_c2":
        fc = G.real()
        if F is None:
            L.orc_bcut_calc_c_f_c2(P(fc), P(x), *[P(a) for a in v], cpp, ipp, P(pid), D(dx), D(dt), D(0.0),
                                   s3, I(g))
        else:
            F.bcut_calc_c_f_c2_(P(fc), P(x), *[P(a) for a in v], *cp, *ip_, P(pid), _r(dx), _r(dt), _r(0.0),
                                s3, _i(g))
        out["fc"] = fc
    elif name == "calc_d_u":
        ud = G.real()
        if F is None:
            L.orc_bcut_calc_d_u(P(ud), P(x), cpp, ipp, P(pid), D(rho), D(mu), D(dx), D(dt), D(0.3), s3, I(g))
        else:
            F.bcut_calc_d_u_(P(ud), P(x), *cp, *ip_, P(pid), _r(rho), _r(mu), _r(dx), _r(dt), _r(0.3), s3, _i(g))
        out["ud"] = ud
    elif name == "update_u":
        uc, ucp, ud, udp = G.real(), G.real(), G.real(), G.real()
        for d in range(3):
            u = x.copy()
            if F is None:
                L.orc_bcut_update_u(P(u), P(uc), P(ucp), P(ud), P(udp), P(y), cpp, ipp, P(pid), I(d), D(rho),
                                    D(mu), D(dx), D(dt), D(0.2), D(0.1), D(-0.2), D(0.3), s3, I(g))
            else:
                F.bcut_update_u_(P(u), P(uc), P(ucp), P(ud), P(udp), P(y), *cp, *ip_, P(pid), _i(d), _r(rho),
                                 _r(mu), _r(dx), _r(dt), _r(0.2), _r(0.1), _r(-0.2), _r(0.3), s3, _i(g))
            out["u%d" % d] = u
    elif name == "calc_abd_u":
        uc, ucp, p = G.real(), G.real(), G.real()
        for d in range(3):
            Aout = [G.real() for _ in range(7)]
            bout, u, ud = G.real(), x.copy(), G.real()
            if F is None:
                L.orc_bcut_calc_abd_u(O.parr(Aout), P(bout), P(u), P(uc), P(ucp), P(ud), P(p), cpp, ipp, P(pid),
                                      I(d), D(0.6), D(rho), D(mu), D(dx), D(dt), D(0.1), D(0.0), D(0.0), D(0.0),
                                      s3, I(g))
            else:
                F.bcut_calc_abd_u_(*[P(a) for a in Aout], P(bout), P(u), P(uc), P(ucp), P(ud), P(p), *cp, *ip_,
                                   P(pid), _i(d), _r(0.6), _r(rho), _r(mu), _r(dx), _r(dt), _r(0.1), _r(0.0),
                                   _r(0.0), _r(0.0), s3, _i(g))
            for q in range(7):
                out["A%d_%d" % (q, d)] = Aout[q]
            out["b_%d" % d] = bout
            out["u_%d" % d] = u
            out["ud_%d" % d] = ud
    elif name == "add_g_f":
        ux, uy, uz, t = x.copy(), y.copy(), b.copy(), G.real()
        fx, fy, fz = G.real(), G.real(), G.real()
        if F is None:
            L.orc_bcut_add_g(P(ux), P(uy), P(uz), P(t), D(0.1), D(-0.2), D(-9.8), D(0.03), D(0.5), D(dx), D(dt),
                             s3, I(g))
            L.orc_bcut_add_f(P(ux), P(uy), P(uz), P(fx), P(fy), P(fz), P(pid), D(rho), D(dx), D(dt), s3, I(g))
        else:
            F.bcut_add_g_(P(ux), P(uy), P(uz), P(t), _r(0.1), _r(-0.2), _r(-9.8), _r(0.03), _r(0.5), _r(dx),
                          _r(dt), s3, _i(g))
            F.bcut_add_f_(P(ux), P(uy), P(uz), P(fx), P(fy), P(fz), P(pid), _r(rho), _r(dx), _r(dt), s3, _i(g))
        out.update(ux=ux, uy=uy, uz=uz)
    elif name == "remove_p":
        ux, uy, uz = x.copy(), y.copy(), b.copy()
        p = G.real()
        if F is None:
            L.orc_bcut_remove_p(P(ux), P(uy), P(uz), P(p), cpp, ipp, P(pid), D(rho), D(dx), D(dt), s3, I(g))
        else:
            F.bcut_remove_p_(P(ux), P(uy), P(uz), P(p), *cp, *ip_, P(pid), _r(rho), _r(dx), _r(dt), s3, _i(g))
        out.update(ux=ux, uy=uy, uz=uz)
    elif name == "calc_ab_p":
        Aout = [G.real() for _ in range(7)]
        bout, p = G.real(), G.real()
        ux, uy, uz = x.copy(), y.copy(), b.copy()
        if F is None:
            L.orc_bcut_calc_ab_p(O.parr(Aout), P(bout), O.parr(v), P(ux), P(uy), P(uz), cpp, ipp, P(pid), D(rho),
                                 D(dx), D(dt), s3, I(g))
        else:
            F.bcut_calc_ab_p_(*[P(a) for a in Aout], P(bout), *[P(a) for a in v], P(p), P(ux), P(uy), P(uz),
                              *cp, *ip_, P(pid), _r(rho), _r(dx), _r(dt), s3, _i(g))
        for q in range(7):
            out["A%d" % q] = Aout[q]
        out["b"] = bout
        for q in range(6):
            out["v%d" % q] = v[q]
    elif name == "corr_u":
        ux, uy, uz, lapp, p = x.copy(), y.copy(), b.copy(), G.real(), G.real()
        if F is None:
            L.orc_bcut_corr_u(P(ux), P(uy), P(uz), O.parr(v), P(lapp), P(p), cpp, ipp, P(pid), D(rho), D(dx),
                              D(dt), s3, I(g))
        else:
            F.bcut_corr_u_(P(ux), P(uy), P(uz), *[P(a) for a in v], P(lapp), P(p), *cp, *ip_, P(pid), _r(rho),
                           _r(dx), _r(dt), s3, _i(g))
        out.update(ux=ux, uy=uy, uz=uz, lapp=lapp)
        for q in range(6):
            out["v%d" % q] = v[q]
    elif name == "set_reference_value":
        Aout = [a.copy() for a in A]
        bout = b.copy()
        org = np.array([0.25, -0.5, 1.0])
        sz = [s3[0], s3[1], s3[2]]
        xr = org[0] + dx * (sz[0] // 2 + 0.37)
        yr = org[1] + dx * (sz[1] // 2 + 0.61)
        zr = org[2] + dx * (sz[2] // 2 + 0.13)
        if F is None:
            L.orc_bcut_set_reference_value(O.parr(Aout), P(bout), D(xr), D(yr), D(zr), D(0.4), D(dx), P(org),
                                           s3, I(g))
        else:
            F.bcut_set_reference_value_(*[P(a) for a in Aout], P(bout), _r(xr), _r(yr), _r(zr), _r(0.4), _r(dx),
                                        P(org), s3, _i(g))
        for q in range(7):
            out["A%d" % q] = Aout[q]
        out["b"] = bout
    elif name == "calc_f_p":
        fx, fy, fz = G.real(), G.real(), G.real()
        fsp = np.zeros(3)
        if F is None:
            L.orc_bcut_calc_f_p(P(fx), P(fy), P(fz), P(fsp), I(1), P(x), cpp, ipp, P(pid), D(dx), D(dt), s3, I(g))
        else:
            F.bcut_calc_f_p_(P(fx), P(fy), P(fz), P(fsp), _i(1), P(x), *cp, *ip_, P(pid), _r(dx), _r(dt), s3,
                             _i(g))
        out.update(fx=fx, fy=fy, fz=fz, fsp=fsp)
    elif name == "calc_f_v":
        ux, uy, uz = G.real(), G.real(), G.real()
        fx, fy, fz = G.real(), G.real(), G.real()
        fsv = np.zeros(3)
        if F is None:
            L.orc_bcut_calc_f_v(P(fx), P(fy), P(fz), P(fsv), I(1), P(ux), P(uy), P(uz), cpp, ipp, P(pid), D(1.0),
                                D(0.01), D(dx), D(dt), D(0.0), s3, I(g))
        else:
            F.bcut_calc_f_v_(P(fx), P(fy), P(fz), P(fsv), _i(1), P(ux), P(uy), P(uz), *cp, *ip_, P(pid), _r(1.0),
                             _r(0.01), _r(dx), _r(dt), _r(0.0), s3, _i(g))
        out.update(fx=fx, fy=fy, fz=fz, fsv=fsv)
    elif name == "bstl":
        cc = [a.copy() for a in c]
        ii = [a.copy() for a in cid]
        # a few exact zero cuts for detect_zerocut
        cc[2][tuple(s // 2 for s in cc[2].shape)] = 0.0
        ii[2][tuple(s // 2 for s in cc[2].shape)] = 2
        if F is None:
            L.orc_bstl_cutoff(O.parr(cc), O.parr(ii, ctypes.c_int), D(0.0), s3, I(g))
            nz = L.orc_bstl_detect_zerocut(O.parr(cc), O.parr(ii, ctypes.c_int), s3, I(g))
            L.orc_bstl_cutoff(O.parr(cc), O.parr(ii, ctypes.c_int), D(0.05), s3, I(g))
            n1 = L.orc_bstl_fill_holes_v1(O.parr(cc), O.parr(ii, ctypes.c_int), I(1), s3, I(g))
            n2 = L.orc_bstl_fill_holes_v2(O.parr(cc), O.parr(ii, ctypes.c_int), I(1), s3, I(g))
        else:
            n = I(0)
            F.bstl_cutoff_(*[P(a) for a in cc], *[P(a) for a in ii], _r(0.0), s3, _i(g))
            F.bstl_detect_zerocut_(*[P(a) for a in cc], *[P(a) for a in ii], ctypes.byref(n), s3, _i(g))
            nz = n.value
            F.bstl_cutoff_(*[P(a) for a in cc], *[P(a) for a in ii], _r(0.05), s3, _i(g))
            F.bstl_fill_holes_v1_(*[P(a) for a in cc], *[P(a) for a in ii], ctypes.byref(n), _i(1), s3, _i(g))
            n1 = n.value
            F.bstl_fill_holes_v2_(*[P(a) for a in cc], *[P(a) for a in ii], ctypes.byref(n), _i(1), s3, _i(g))
            n2 = n.value
        for q in range(6):
            out["c%d" % q] = cc[q]
            out["cid%d" % q] = ii[q]
        out["counts"] = np.array([nz, n1, n2], dtype=np.int64)
    else:
        raise KeyError(name)
    return out


CASES = ["jacobi_smoother", "rbgs_smoother", "calc_ax", "calc_r", "calc_r2", "blas", "sup_copy", "sup_c2f",
         "sup_f2c", "bc_x", "bc_a", "calc_c_f_c2", "calc_d_u", "update_u", "calc_abd_u", "add_g_f", "remove_p",
         "calc_ab_p", "corr_u", "set_reference_value", "calc_f_p", "calc_f_v", "bstl"]
