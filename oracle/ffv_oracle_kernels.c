/*
 * ffv_oracle_kernels.c -- per-block kernels of the CPU oracle (test infrastructure only).
 *
 * Each function restates one reference Fortran subroutine in plain C with the
 * same arithmetic in the same order (Fortran evaluates a*b/c as (a*b)/c and
 * a-b-c as (a-b)-c; the C below spells that out).  Compiled with
 * -ffp-contract=off so no FMA changes a rounding.  The tests compare these
 * bit-for-bit with the reference sources compiled all-double.
 */
#include "ffv_oracle.h"

#include <math.h>
#include <stddef.h>

/* Fortran index (i,j,k), each in [1-g, n+g], into the padded block. */
#define DECL_SZ                             \
  const int ix = sz[0], jx = sz[1], kx = sz[2]; \
  const ptrdiff_t nxp = ix + 2 * g, nyp = jx + 2 * g; \
  (void)ix; (void)jx; (void)kx;
#define IX(i, j, k) ((ptrdiff_t)((i) - 1 + g) + nxp * ((ptrdiff_t)((j) - 1 + g) + nyp * ((ptrdiff_t)((k) - 1 + g))))

/* ------------------------------------------------------------------ bils.f90 */

/* bils.f90:5-55 jacobi_smoother */
void orc_jacobi_smoother(real* x1, const real* x0, const real* Ap, const real* Aw, const real* Ae,
                         const real* As, const real* An, const real* Ab, const real* At, const real* b,
                         real omega, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real r = b[m] - Aw[m] * x0[IX(i - 1, j, k)] - Ae[m] * x0[IX(i + 1, j, k)] -
                 As[m] * x0[IX(i, j - 1, k)] - An[m] * x0[IX(i, j + 1, k)] -
                 Ab[m] * x0[IX(i, j, k - 1)] - At[m] * x0[IX(i, j, k + 1)];
        real x_tmp = r / Ap[m];
        x1[m] = x0[m] - omega * (x0[m] - x_tmp);
      }
}

/* bils.f90:57-108 rbgs_smoother; colour parity mod(k+j+color+offset,2) on block-local
 * Fortran indices (bils.f90:88). */
void orc_rbgs_smoother(real* x, const real* Ap, const real* Aw, const real* Ae, const real* As,
                       const real* An, const real* Ab, const real* At, const real* b, real omega,
                       int color, int offset, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1 + (k + j + color + offset) % 2; i <= ix; i += 2) {
        ptrdiff_t m = IX(i, j, k);
        real r = b[m] - Aw[m] * x[IX(i - 1, j, k)] - Ae[m] * x[IX(i + 1, j, k)] -
                 As[m] * x[IX(i, j - 1, k)] - An[m] * x[IX(i, j + 1, k)] -
                 Ab[m] * x[IX(i, j, k - 1)] - At[m] * x[IX(i, j, k + 1)];
        real x_tmp = r / Ap[m];
        x[m] = x[m] - omega * (x[m] - x_tmp);
      }
}

/* bils.f90:162-206 calc_ax */
void orc_calc_ax(real* Ax, const real* Ap, const real* Aw, const real* Ae, const real* As,
                 const real* An, const real* Ab, const real* At, const real* x, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        Ax[m] = Aw[m] * x[IX(i - 1, j, k)] + Ae[m] * x[IX(i + 1, j, k)] + As[m] * x[IX(i, j - 1, k)] +
                An[m] * x[IX(i, j + 1, k)] + Ab[m] * x[IX(i, j, k - 1)] + At[m] * x[IX(i, j, k + 1)] +
                Ap[m] * x[m];
      }
}

/* bils.f90:208-254 calc_r */
void orc_calc_r(real* r, const real* Ap, const real* Aw, const real* Ae, const real* As,
                const real* An, const real* Ab, const real* At, const real* x, const real* b,
                const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        r[m] = b[m] - Aw[m] * x[IX(i - 1, j, k)] - Ae[m] * x[IX(i + 1, j, k)] -
               As[m] * x[IX(i, j - 1, k)] - An[m] * x[IX(i, j + 1, k)] -
               Ab[m] * x[IX(i, j, k - 1)] - At[m] * x[IX(i, j, k + 1)] - Ap[m] * x[m];
      }
}

/* bils.f90:256-304 calc_r2 */
real orc_calc_r2(const real* Ap, const real* Aw, const real* Ae, const real* As, const real* An,
                 const real* Ab, const real* At, const real* x, const real* b, const int* sz, int g) {
  DECL_SZ
  real rr = 0.0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real r = b[m] - Aw[m] * x[IX(i - 1, j, k)] - Ae[m] * x[IX(i + 1, j, k)] -
                 As[m] * x[IX(i, j - 1, k)] - An[m] * x[IX(i, j + 1, k)] -
                 Ab[m] * x[IX(i, j, k - 1)] - At[m] * x[IX(i, j, k + 1)] - Ap[m] * x[m];
        rr = rr + r * r;
      }
  return rr;
}

/* ------------------------------------------------------------------ blas.f90 */

/* blas.f90:5-33 fill (whole padded block) */
void orc_fill(real* x, real a, const int* sz, int g) {
  DECL_SZ
  for (int k = 1 - g; k <= kx + g; k++)
    for (int j = 1 - g; j <= jx + g; j++)
      for (int i = 1 - g; i <= ix + g; i++) x[IX(i, j, k)] = a;
}

/* blas.f90:69-97 copy (whole padded block) */
void orc_copy(real* y, const real* x, const int* sz, int g) {
  DECL_SZ
  for (int k = 1 - g; k <= kx + g; k++)
    for (int j = 1 - g; j <= jx + g; j++)
      for (int i = 1 - g; i <= ix + g; i++) y[IX(i, j, k)] = x[IX(i, j, k)];
}

/* blas.f90:162-193 adda (interior) */
void orc_adda(real* y, real a, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) y[IX(i, j, k)] = y[IX(i, j, k)] + a;
}

/* blas.f90:329-361 axpy */
void orc_axpy(real* y, const real* x, real a, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        y[m] = y[m] + a * x[m];
      }
}

/* blas.f90:363-395 xpay */
void orc_xpay(real* y, const real* x, real a, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        y[m] = x[m] + a * y[m];
      }
}

/* blas.f90:397-430 axpyz */
void orc_axpyz(real* z, const real* x, const real* y, real a, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        z[m] = a * x[m] + y[m];
      }
}

/* blas.f90:432-464 axpbypz */
void orc_axpbypz(real* z, const real* x, const real* y, real a, real b, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        z[m] = a * x[m] + b * y[m] + z[m];
      }
}

/* blas.f90:466-499 dot */
real orc_dot(const real* x, const real* y, const int* sz, int g) {
  DECL_SZ
  real xy = 0.0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) xy = xy + x[IX(i, j, k)] * y[IX(i, j, k)];
  return xy;
}

/* bsf3d.f90:35-67 sf3d_calc_sum */
real orc_sf3d_calc_sum(const real* x, const int* sz, int g) {
  DECL_SZ
  real s = 0.0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) s = s + x[IX(i, j, k)];
  return s;
}

/* ------------------------------------------------------------------- sup.f90 */

/* sup.f90:18-52 sup_copy_from_neighbor */
void orc_sup_copy_from_neighbor(real* dst, const int* i1_dst, const real* src, const int* i1_src,
                                const int* sz_c, const int* sz, int g) {
  DECL_SZ
  for (int k = 0; k < sz_c[2]; k++)
    for (int j = 0; j < sz_c[1]; j++)
      for (int i = 0; i < sz_c[0]; i++)
        dst[IX(i + i1_dst[0], j + i1_dst[1], k + i1_dst[2])] =
            src[IX(i + i1_src[0], j + i1_src[1], k + i1_src[2])];
}

/* sup.f90:1-16 sup_get_intp_index */
static void intp_index(int* ic, real* r, int i, int n) {
  *ic = i / 2;
  *r = 0.25 + 0.5 * (real)(i % 2);
  if (i < 2) {
    *ic = 1;
    *r = -0.25;
  } else if (i > 2 * n - 1) {
    *ic = n - 1;
    *r = 1.25;
  }
}

/* sup.f90:54-120 sup_copy_from_neighbor_c2f (trilinear from the coarse interior) */
void orc_sup_copy_from_neighbor_c2f(real* dst, const int* i1_dst, const real* src, const int* i1_src,
                                    const int* sz_c, const int* sz, int g) {
  DECL_SZ
  for (int k = 0; k < sz_c[2]; k++)
    for (int j = 0; j < sz_c[1]; j++)
      for (int i = 0; i < sz_c[0]; i++) {
        int i1, j1, k1;
        real r, s, t;
        intp_index(&i1, &r, i + i1_src[0], ix);
        intp_index(&j1, &s, j + i1_src[1], jx);
        intp_index(&k1, &t, k + i1_src[2], kx);
        real nnn = src[IX(i1, j1, k1)], pnn = src[IX(i1 + 1, j1, k1)];
        real npn = src[IX(i1, j1 + 1, k1)], nnp = src[IX(i1, j1, k1 + 1)];
        real npp = src[IX(i1, j1 + 1, k1 + 1)], pnp = src[IX(i1 + 1, j1, k1 + 1)];
        real ppn = src[IX(i1 + 1, j1 + 1, k1)], ppp = src[IX(i1 + 1, j1 + 1, k1 + 1)];
        real a0 = (1.0 - r) * nnn + r * pnn;
        real a1 = (1.0 - r) * npn + r * ppn;
        real a2 = (1.0 - r) * nnp + r * pnp;
        real a3 = (1.0 - r) * npp + r * ppp;
        real c0 = (1.0 - s) * a0 + s * a1;
        real c1 = (1.0 - s) * a2 + s * a3;
        dst[IX(i + i1_dst[0], j + i1_dst[1], k + i1_dst[2])] = (1.0 - t) * c0 + t * c1;
      }
}

/* sup.f90:122-179 sup_copy_from_neighbor_f2c (mean of 8 fine cells, sum order nnn pnn npn nnp
 * npp pnp ppn ppp) */
void orc_sup_copy_from_neighbor_f2c(real* dst, const int* i1_dst, const real* src, const int* i1_src,
                                    const int* sz_c, const int* sz, int g) {
  DECL_SZ
  for (int k = 0; k < sz_c[2]; k++)
    for (int j = 0; j < sz_c[1]; j++)
      for (int i = 0; i < sz_c[0]; i++) {
        int i1 = 2 * (i + i1_src[0]) - 1, j1 = 2 * (j + i1_src[1]) - 1, k1 = 2 * (k + i1_src[2]) - 1;
        real v = src[IX(i1, j1, k1)] + src[IX(i1 + 1, j1, k1)] + src[IX(i1, j1 + 1, k1)] +
                 src[IX(i1, j1, k1 + 1)] + src[IX(i1, j1 + 1, k1 + 1)] + src[IX(i1 + 1, j1, k1 + 1)] +
                 src[IX(i1 + 1, j1 + 1, k1)] + src[IX(i1 + 1, j1 + 1, k1 + 1)];
        dst[IX(i + i1_dst[0], j + i1_dst[1], k + i1_dst[2])] = 0.125 * v;
      }
}

/* ------------------------------------------------------ Scalar3DUpdater5.h */
/* 0-based (C++ Index3DS) access used by the updater: (i,j,k) in [-g, n+g) */
#define IX0(i, j, k) IX((i) + 1, (j) + 1, (k) + 1)

/* Scalar3DUpdater5.h:198-203 interpolateF2C: sum order f(i,j,k) f(i+1,j,k) f(i,j+1,k)
 * f(i+1,j+1,k) f(i,j,k+1) f(i+1,j,k+1) f(i,j+1,k+1) f(i+1,j+1,k+1) */
static real u5_f2c_val(const real* f, int I, int J, int K, const int* sz, int g) {
  DECL_SZ
  int i = 2 * I, j = 2 * J, k = 2 * K;
  return 0.125 * (f[IX0(i, j, k)] + f[IX0(i + 1, j, k)] + f[IX0(i, j + 1, k)] + f[IX0(i + 1, j + 1, k)] +
                  f[IX0(i, j, k + 1)] + f[IX0(i + 1, j, k + 1)] + f[IX0(i, j + 1, k + 1)] +
                  f[IX0(i + 1, j + 1, k + 1)]);
}

/* Scalar3DUpdater5.h:650-763 copyFromNeighborF2C_0: coarse ghosts on `face` of the coarse block
 * from the fine neighbour that sits at `subface`. */
void orc_u5_f2c(real* c, const real* f, int face, int subface, const int* sz, int g) {
  DECL_SZ
  const int nx = ix, ny = jx, nz = kx, vc = g;
  const int o0 = subface & 1, o1 = subface >> 1;
  switch (face) {
    case 0: {
      int j0 = (ny / 2) * o0, k0 = (nz / 2) * o1;
      for (int k = 0; k < nz / 2; k++)
        for (int j = 0; j < ny / 2; j++)
          for (int i = 0; i < vc; i++)
            c[IX0(i - vc, j + j0, k + k0)] = u5_f2c_val(f, i + nx / 2 - vc, j, k, sz, g);
      break;
    }
    case 1: {
      int j0 = (ny / 2) * o0, k0 = (nz / 2) * o1;
      for (int k = 0; k < nz / 2; k++)
        for (int j = 0; j < ny / 2; j++)
          for (int i = 0; i < vc; i++) c[IX0(i + nx, j + j0, k + k0)] = u5_f2c_val(f, i, j, k, sz, g);
      break;
    }
    case 2: {
      int k0 = (nz / 2) * o0, i0 = (nx / 2) * o1;
      for (int k = 0; k < nz / 2; k++)
        for (int j = 0; j < vc; j++)
          for (int i = 0; i < nx / 2; i++)
            c[IX0(i + i0, j - vc, k + k0)] = u5_f2c_val(f, i, j + ny / 2 - vc, k, sz, g);
      break;
    }
    case 3: {
      int k0 = (nz / 2) * o0, i0 = (nx / 2) * o1;
      for (int k = 0; k < nz / 2; k++)
        for (int j = 0; j < vc; j++)
          for (int i = 0; i < nx / 2; i++) c[IX0(i + i0, j + ny, k + k0)] = u5_f2c_val(f, i, j, k, sz, g);
      break;
    }
    case 4: {
      int i0 = (nx / 2) * o0, j0 = (ny / 2) * o1;
      for (int k = 0; k < vc; k++)
        for (int j = 0; j < ny / 2; j++)
          for (int i = 0; i < nx / 2; i++)
            c[IX0(i + i0, j + j0, k - vc)] = u5_f2c_val(f, i, j, k + nz / 2 - vc, sz, g);
      break;
    }
    case 5: {
      int i0 = (nx / 2) * o0, j0 = (ny / 2) * o1;
      for (int k = 0; k < vc; k++)
        for (int j = 0; j < ny / 2; j++)
          for (int i = 0; i < nx / 2; i++) c[IX0(i + i0, j + j0, k + nz)] = u5_f2c_val(f, i, j, k, sz, g);
      break;
    }
  }
}

/* Scalar3DUpdater5.h:765-885 copyFromNeighborC2F_0: f_ghost = f_adjacent + 0.5*(c_in - c_ghost)
 * (CalcGradient 245-252, Gradient 255-259); layers filled in order, the outer layer steps from
 * the inner one. */
void orc_u5_c2f(real* f, const real* c, int face, int subface, const int* sz, int g) {
  DECL_SZ
  const int nx = ix, ny = jx, nz = kx, vc = g;
  const int o0 = subface & 1, o1 = subface >> 1;
  switch (face) {
    case 0: {
      int J0 = ny * o0, K0 = nz * o1;
      for (int K = 0; K < nz; K++)
        for (int J = 0; J < ny; J++)
          for (int I = 0; I < vc; I++) {
            int cj = J / 2 + J0 / 2, ck = K / 2 + K0 / 2;
            real dp = (c[IX0(nx - 1, cj, ck)] - c[IX0(nx, cj, ck)]) * 0.5;
            f[IX0(-1 - I, J, K)] = f[IX0(0 - I, J, K)] + dp;
          }
      break;
    }
    case 1: {
      int J0 = ny * o0, K0 = nz * o1;
      for (int K = 0; K < nz; K++)
        for (int J = 0; J < ny; J++)
          for (int I = 0; I < vc; I++) {
            int cj = J / 2 + J0 / 2, ck = K / 2 + K0 / 2;
            real dp = (c[IX0(0, cj, ck)] - c[IX0(-1, cj, ck)]) * 0.5;
            f[IX0(nx + I, J, K)] = f[IX0(nx - 1 + I, J, K)] + dp;
          }
      break;
    }
    case 2: {
      int K0 = nz * o0, I0 = nx * o1;
      for (int K = 0; K < nz; K++)
        for (int J = 0; J < vc; J++)
          for (int I = 0; I < nx; I++) {
            int ci = I / 2 + I0 / 2, ck = K / 2 + K0 / 2;
            real dp = (c[IX0(ci, ny - 1, ck)] - c[IX0(ci, ny, ck)]) * 0.5;
            f[IX0(I, -1 - J, K)] = f[IX0(I, 0 - J, K)] + dp;
          }
      break;
    }
    case 3: {
      int K0 = nz * o0, I0 = nx * o1;
      for (int K = 0; K < nz; K++)
        for (int J = 0; J < vc; J++)
          for (int I = 0; I < nx; I++) {
            int ci = I / 2 + I0 / 2, ck = K / 2 + K0 / 2;
            real dp = (c[IX0(ci, 0, ck)] - c[IX0(ci, -1, ck)]) * 0.5;
            f[IX0(I, ny + J, K)] = f[IX0(I, ny - 1 + J, K)] + dp;
          }
      break;
    }
    case 4: {
      int I0 = nx * o0, J0 = ny * o1;
      for (int K = 0; K < vc; K++)
        for (int J = 0; J < ny; J++)
          for (int I = 0; I < nx; I++) {
            int ci = I / 2 + I0 / 2, cj = J / 2 + J0 / 2;
            real dp = (c[IX0(ci, cj, nz - 1)] - c[IX0(ci, cj, nz)]) * 0.5;
            f[IX0(I, J, -1 - K)] = f[IX0(I, J, 0 - K)] + dp;
          }
      break;
    }
    case 5: {
      int I0 = nx * o0, J0 = ny * o1;
      for (int K = 0; K < vc; K++)
        for (int J = 0; J < ny; J++)
          for (int I = 0; I < nx; I++) {
            int ci = I / 2 + I0 / 2, cj = J / 2 + J0 / 2;
            real dp = (c[IX0(ci, cj, 0)] - c[IX0(ci, cj, -1)]) * 0.5;
            f[IX0(I, J, nz + K)] = f[IX0(I, J, nz - 1 + K)] + dp;
          }
      break;
    }
  }
}

/* ------------------------------------------------------------------- bcx.f90 */
/* Face -> routine map BlockScalar3D.h:201-254: X_M x3, X_P x1, Y_M x4, Y_P x2, Z_M x6, Z_P x5.
 * Dirichlet bcx.f90:5-147, Neumann 294-436, periodic 585-658 (the minus-side periodic is a
 * no-op, BlockScalar3D.h:207-209).  Ghost planes span the whole padded face. */
void orc_bc_x(real* x, int face, int type, real xc, const int* sz, int g) {
  DECL_SZ
  const int axis = face / 2, plus = face & 1;
  const ptrdiff_t stride[3] = {1, nxp, nxp * nyp};
  const int o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
  const int n = sz[axis];
  for (int c2 = 1 - g; c2 <= sz[o2] + g; c2++)
    for (int c1 = 1 - g; c1 <= sz[o1] + g; c1++) {
      /* element with normal index q: base + (q - 1 + g) * stride[axis] */
      ptrdiff_t base = (ptrdiff_t)(c1 - 1 + g) * stride[o1] + (ptrdiff_t)(c2 - 1 + g) * stride[o2];
#define AT(q) x[base + (ptrdiff_t)((q) - 1 + g) * stride[axis]]
      if (type == 0) {
        if (plus) {
          AT(n + 1) = 2.0 * xc - AT(n);
          AT(n + 2) = xc;
        } else {
          AT(0) = 2.0 * xc - AT(1);
          AT(-1) = xc;
        }
      } else if (type == 1) {
        if (plus) {
          AT(n + 1) = AT(n) + xc;
          AT(n + 2) = AT(n) + xc * 2.0;
        } else {
          AT(0) = AT(1) - xc;
          AT(-1) = AT(1) - xc * 2.0;
        }
      } else if (type == 9 && plus) {
        AT(n + 1) = AT(1);
        AT(n + 2) = AT(2);
        AT(0) = AT(n);
        AT(-1) = AT(n - 1);
      }
#undef AT
    }
}

/* bca.f90:5-387 bc_A{w,e,s,n,b,t}_{d,n}: fold the outer-face BC into Ap and b, drop the link. */
void orc_bc_a(real* Ap, real* Anb, real* b, int face, int type, real xc, const int* sz, int g) {
  DECL_SZ
  const int axis = face / 2, plus = face & 1;
  if (type != 0 && type != 1) return;
  int lo[3] = {1, 1, 1}, hi[3] = {ix, jx, kx};
  lo[axis] = hi[axis] = plus ? sz[axis] : 1;
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        ptrdiff_t m = IX(i, j, k);
        if (type == 0) {
          Ap[m] = Ap[m] - Anb[m];
          b[m] = b[m] - Anb[m] * xc * 2.0;
        } else {
          /* ghost = x -/+ xc on the minus/plus side (bcx.f90:294-436) */
          Ap[m] = Ap[m] + Anb[m];
          b[m] = plus ? b[m] - Anb[m] * xc : b[m] + Anb[m] * xc;
        }
        Anb[m] = 0.0;
      }
}

/* ------------------------------------------------------------------ bcut.f90 */

#define LOAD_GEOM(m)                                                                    \
  const real d0 = c[0][m], d1 = c[1][m], d2 = c[2][m], d3 = c[3][m], d4 = c[4][m], d5 = c[5][m]; \
  const int cidp0 = cid[0][m], cidp1 = cid[1][m], cidp2 = cid[2][m], cidp3 = cid[3][m],          \
            cidp4 = cid[4][m], cidp5 = cid[5][m];                                                \
  const int pidp = pid[m];

/* bcut.f90:1583-1764 bcut_calc_c_f_c2: flux-form 2nd-order convection. */
void orc_bcut_calc_c_f_c2(real* fc, const real* f, const real* vw, const real* ve, const real* vs,
                          const real* vn, const real* vb, const real* vt, const real* const* c,
                          const int* const* cid, const int* pid, real dx, real dt, real fi,
                          const int* sz, int g) {
  DECL_SZ
  (void)dt; (void)fi;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real vx0 = vw[m], vx1 = ve[m], vy2 = vs[m], vy3 = vn[m], vz4 = vb[m], vz5 = vt[m];
        real fp = f[m], fw = f[IX(i - 1, j, k)], fe = f[IX(i + 1, j, k)], fs = f[IX(i, j - 1, k)],
             fn = f[IX(i, j + 1, k)], fb = f[IX(i, j, k - 1)], ft = f[IX(i, j, k + 1)];
        LOAD_GEOM(m)
        real f0 = 0.5 * (fp + fw), f1 = 0.5 * (fp + fe), f2 = 0.5 * (fp + fs);
        real f3 = 0.5 * (fp + fn), f4 = 0.5 * (fp + fb), f5 = 0.5 * (fp + ft);
        real q0 = vx0 * f0, q1 = vx1 * f1, q2 = vy2 * f2, q3 = vy3 * f3, q4 = vz4 * f4, q5 = vz5 * f5;
        if (cidp0 != 0) q0 = (d0 - 0.5) / (d0 + 0.5) * q1;
        if (cidp1 != 0) q1 = (d1 - 0.5) / (d1 + 0.5) * q0;
        if (cidp2 != 0) q2 = (d2 - 0.5) / (d2 + 0.5) * q3;
        if (cidp3 != 0) q3 = (d3 - 0.5) / (d3 + 0.5) * q2;
        if (cidp4 != 0) q4 = (d4 - 0.5) / (d4 + 0.5) * q5;
        if (cidp5 != 0) q5 = (d5 - 0.5) / (d5 + 0.5) * q4;
        if (cidp0 != 0 && cidp1 != 0) { q0 = 0.0; q1 = 0.0; }
        if (cidp2 != 0 && cidp3 != 0) { q2 = 0.0; q3 = 0.0; }
        if (cidp4 != 0 && cidp5 != 0) { q4 = 0.0; q5 = 0.0; }
        fc[m] = (q1 - q0) / dx + (q3 - q2) / dx + (q5 - q4) / dx;
        if (pidp != 1) fc[m] = 0.0;
      }
}

/* bcut.f90:3397-3552 bcut_calc_d_u: diffusion with cut-face viscosity mu/d*2/(d0+d1), wall Uc. */
void orc_bcut_calc_d_u(real* ud, const real* u, const real* const* c, const int* const* cid,
                       const int* pid, real rho, real mu, real dx, real dt, real Uc, const int* sz, int g) {
  DECL_SZ
  (void)dt;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real mu0 = mu, mu1 = mu, mu2 = mu, mu3 = mu, mu4 = mu, mu5 = mu;
        real m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, m4 = 0.0, m5 = 0.0;
        LOAD_GEOM(m)
        if (cidp0 != 0) { mu0 = mu / d0 * 2.0 / (d0 + d1); mu1 = mu / d1 * 2.0 / (d0 + d1); m0 = 1.0; }
        if (cidp1 != 0) { mu0 = mu / d0 * 2.0 / (d0 + d1); mu1 = mu / d1 * 2.0 / (d0 + d1); m1 = 1.0; }
        if (cidp2 != 0) { mu2 = mu / d2 * 2.0 / (d2 + d3); mu3 = mu / d3 * 2.0 / (d2 + d3); m2 = 1.0; }
        if (cidp3 != 0) { mu2 = mu / d2 * 2.0 / (d2 + d3); mu3 = mu / d3 * 2.0 / (d2 + d3); m3 = 1.0; }
        if (cidp4 != 0) { mu4 = mu / d4 * 2.0 / (d4 + d5); mu5 = mu / d5 * 2.0 / (d4 + d5); m4 = 1.0; }
        if (cidp5 != 0) { mu4 = mu / d4 * 2.0 / (d4 + d5); mu5 = mu / d5 * 2.0 / (d4 + d5); m5 = 1.0; }
        real l0 = mu0 / rho / (dx * dx), l1 = mu1 / rho / (dx * dx), l2 = mu2 / rho / (dx * dx);
        real l3 = mu3 / rho / (dx * dx), l4 = mu4 / rho / (dx * dx), l5 = mu5 / rho / (dx * dx);
        real up = u[m];
        real uw = u[IX(i - 1, j, k)] * (1.0 - m0) + Uc * m0;
        real ue = u[IX(i + 1, j, k)] * (1.0 - m1) + Uc * m1;
        real us = u[IX(i, j - 1, k)] * (1.0 - m2) + Uc * m2;
        real un = u[IX(i, j + 1, k)] * (1.0 - m3) + Uc * m3;
        real ub = u[IX(i, j, k - 1)] * (1.0 - m4) + Uc * m4;
        real ut = u[IX(i, j, k + 1)] * (1.0 - m5) + Uc * m5;
        ud[m] = l1 * (ue - up) - l0 * (up - uw) + l3 * (un - up) - l2 * (up - us) + l5 * (ut - up) -
                l4 * (up - ub);
        if (pidp != 1) ud[m] = 0.0;
      }
}

/* Pressure-gradient face terms shared by update_u (4299-4523), calc_abd_u (43-322), remove_p
 * (3879-4061): dp on the six faces, mirrored with (d-0.5)/(d+0.5) at a cut face, zero if both
 * faces along an axis are cut.  `gterm` reproduces the "- g*0.0" terms of update_u/abd_u. */
typedef struct { real x0, x1, y2, y3, z4, z5; } face6;
static inline face6 grad_faces(const real* p, ptrdiff_t m, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz_,
                               real r0, real dx, int with_g, real gx, real gy, real gz, real d0, real d1,
                               real d2, real d3, real d4, real d5, int cidp0, int cidp1, int cidp2,
                               int cidp3, int cidp4, int cidp5) {
  real dpx0 = (p[m] - p[m - sx]) / dx, dpx1 = (p[m + sx] - p[m]) / dx;
  real dpy2 = (p[m] - p[m - sy]) / dx, dpy3 = (p[m + sy] - p[m]) / dx;
  real dpz4 = (p[m] - p[m - sz_]) / dx, dpz5 = (p[m + sz_] - p[m]) / dx;
  face6 r;
  if (with_g) {
    r.x0 = r0 * dpx0 - gx * 0.0; r.x1 = r0 * dpx1 - gx * 0.0;
    r.y2 = r0 * dpy2 - gy * 0.0; r.y3 = r0 * dpy3 - gy * 0.0;
    r.z4 = r0 * dpz4 - gz * 0.0; r.z5 = r0 * dpz5 - gz * 0.0;
  } else {
    r.x0 = r0 * dpx0; r.x1 = r0 * dpx1; r.y2 = r0 * dpy2;
    r.y3 = r0 * dpy3; r.z4 = r0 * dpz4; r.z5 = r0 * dpz5;
  }
  if (cidp0 != 0) r.x0 = r.x1 * (d0 - 0.5) / (d0 + 0.5);
  if (cidp1 != 0) r.x1 = r.x0 * (d1 - 0.5) / (d1 + 0.5);
  if (cidp2 != 0) r.y2 = r.y3 * (d2 - 0.5) / (d2 + 0.5);
  if (cidp3 != 0) r.y3 = r.y2 * (d3 - 0.5) / (d3 + 0.5);
  if (cidp4 != 0) r.z4 = r.z5 * (d4 - 0.5) / (d4 + 0.5);
  if (cidp5 != 0) r.z5 = r.z4 * (d5 - 0.5) / (d5 + 0.5);
  if (cidp0 != 0 && cidp1 != 0) { r.x0 = 0.0; r.x1 = 0.0; }
  if (cidp2 != 0 && cidp3 != 0) { r.y2 = 0.0; r.y3 = 0.0; }
  if (cidp4 != 0 && cidp5 != 0) { r.z4 = 0.0; r.z5 = 0.0; }
  return r;
}
#define GRAD_ARGS d0, d1, d2, d3, d4, d5, cidp0, cidp1, cidp2, cidp3, cidp4, cidp5

/* bcut.f90:4299-4523 bcut_update_u (explicit Adams-Bashforth 2, formula 4507-4510) */
void orc_bcut_update_u(real* u, const real* uc, const real* ucp, const real* ud, const real* udp,
                       const real* p, const real* const* c, const int* const* cid, const int* pid,
                       int dir, real rhof, real mu, real dx, real dt, real Uc, real gx, real gy, real gz,
                       const int* sz, int g) {
  DECL_SZ
  (void)mu;
  const ptrdiff_t sy = nxp, sk = nxp * nyp;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        LOAD_GEOM(m)
        face6 r = grad_faces(p, m, 1, sy, sk, 1.0 / rhof, dx, 1, gx, gy, gz, GRAD_ARGS);
        real rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
        real rdp = 0.0;
        if (dir == 0) rdp = rdpx; else if (dir == 1) rdp = rdpy; else if (dir == 2) rdp = rdpz;
        u[m] = u[m] - (1.5 * uc[m] - 0.5 * ucp[m]) * dt + (1.5 * ud[m] - 0.5 * udp[m]) * dt - rdp * dt;
        if (pidp != 1) u[m] = Uc;
      }
}

/* bcut.f90:43-322 bcut_calc_abd_u (implicit U: 7-band matrix and RHS) */
void orc_bcut_calc_abd_u(real* const* A, real* b, real* u, const real* uc, const real* ucp, real* ud,
                         const real* p, const real* const* c, const int* const* cid, const int* pid,
                         int dir, real alpha, real rhof, real mu, real dx, real dt, real Uc,
                         real gx, real gy, real gz, const int* sz, int g) {
  DECL_SZ
  const ptrdiff_t sy = nxp, sk = nxp * nyp;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        LOAD_GEOM(m)
        real mu0 = mu, mu1 = mu, mu2 = mu, mu3 = mu, mu4 = mu, mu5 = mu;
        real m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, m4 = 0.0, m5 = 0.0;
        face6 r = grad_faces(p, m, 1, sy, sk, 1.0 / rhof, dx, 1, gx, gy, gz, GRAD_ARGS);
        if (cidp0 != 0) { mu0 = mu / d0 * 2.0 / (d0 + d1); mu1 = mu / d1 * 2.0 / (d0 + d1); m0 = 1.0; }
        if (cidp1 != 0) { mu0 = mu / d0 * 2.0 / (d0 + d1); mu1 = mu / d1 * 2.0 / (d0 + d1); m1 = 1.0; }
        if (cidp2 != 0) { mu2 = mu / d2 * 2.0 / (d2 + d3); mu3 = mu / d3 * 2.0 / (d2 + d3); m2 = 1.0; }
        if (cidp3 != 0) { mu2 = mu / d2 * 2.0 / (d2 + d3); mu3 = mu / d3 * 2.0 / (d2 + d3); m3 = 1.0; }
        if (cidp4 != 0) { mu4 = mu / d4 * 2.0 / (d4 + d5); mu5 = mu / d5 * 2.0 / (d4 + d5); m4 = 1.0; }
        if (cidp5 != 0) { mu4 = mu / d4 * 2.0 / (d4 + d5); mu5 = mu / d5 * 2.0 / (d4 + d5); m5 = 1.0; }
        if (cidp0 != 0 && cidp1 != 0) { m0 = 1.0; m1 = 1.0; }
        if (cidp2 != 0 && cidp3 != 0) { m2 = 1.0; m3 = 1.0; }
        if (cidp4 != 0 && cidp5 != 0) { m4 = 1.0; m5 = 1.0; }
        real rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
        real rdp = 0.0;
        if (dir == 0) rdp = rdpx; else if (dir == 1) rdp = rdpy; else if (dir == 2) rdp = rdpz;
        real l0 = mu0 / rhof * dt / (dx * dx), l1 = mu1 / rhof * dt / (dx * dx);
        real l2 = mu2 / rhof * dt / (dx * dx), l3 = mu3 / rhof * dt / (dx * dx);
        real l4 = mu4 / rhof * dt / (dx * dx), l5 = mu5 / rhof * dt / (dx * dx);
        real up = u[m];
        real uw = u[IX(i - 1, j, k)] * (1.0 - m0) + Uc * m0;
        real ue = u[IX(i + 1, j, k)] * (1.0 - m1) + Uc * m1;
        real us = u[IX(i, j - 1, k)] * (1.0 - m2) + Uc * m2;
        real un = u[IX(i, j + 1, k)] * (1.0 - m3) + Uc * m3;
        real ub = u[IX(i, j, k - 1)] * (1.0 - m4) + Uc * m4;
        real ut = u[IX(i, j, k + 1)] * (1.0 - m5) + Uc * m5;
        ud[m] = l1 * (ue - up) - l0 * (up - uw) + l3 * (un - up) - l2 * (up - us) + l5 * (ut - up) -
                l4 * (up - ub);
        if (pidp != 1) ud[m] = 0.0;
        A[0][m] = 1.0 + alpha * (l0 + l1) + alpha * (l2 + l3) + alpha * (l4 + l5);
        A[1][m] = -(alpha * l0 * (1.0 - m0));
        A[2][m] = -(alpha * l1 * (1.0 - m1));
        A[3][m] = -(alpha * l2 * (1.0 - m2));
        A[4][m] = -(alpha * l3 * (1.0 - m3));
        A[5][m] = -(alpha * l4 * (1.0 - m4));
        A[6][m] = -(alpha * l5 * (1.0 - m5));
        b[m] = u[m] - (1.5 * uc[m] - 0.5 * ucp[m]) * dt + (1.0 - alpha) * ud[m] * dt - rdp * dt +
               alpha * l0 * Uc * m0 + alpha * l1 * Uc * m1 + alpha * l2 * Uc * m2 + alpha * l3 * Uc * m3 +
               alpha * l4 * Uc * m4 + alpha * l5 * Uc * m5;
        if (pidp != 1) {
          A[0][m] = 1.0;
          for (int q = 1; q < 7; q++) A[q][m] = 0.0;
          b[m] = Uc;
          u[m] = 0.0;
        }
      }
}

/* bcut.f90:3793-3833 bcut_add_g */
void orc_bcut_add_g(real* ux, real* uy, real* uz, const real* t, real gx, real gy, real gz, real betag,
                    real tr, real dx, real dt, const int* sz, int g) {
  DECL_SZ
  (void)dx;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        ux[m] = ux[m] + gx * dt;
        uy[m] = uy[m] + gy * dt;
        uz[m] = uz[m] + gz * dt + betag * (t[m] - tr) * dt;
      }
}

/* bcut.f90:3835-3877 bcut_add_f */
void orc_bcut_add_f(real* ux, real* uy, real* uz, const real* fx, const real* fy, const real* fz,
                    const int* pid, real rhof, real dx, real dt, const int* sz, int g) {
  DECL_SZ
  (void)dx;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        if (pid[m] == 1) {
          ux[m] = ux[m] + fx[m] / rhof * dt;
          uy[m] = uy[m] + fy[m] / rhof * dt;
          uz[m] = uz[m] + fz[m] / rhof * dt;
        }
      }
}

/* bcut.f90:3879-4061 bcut_remove_p: add back the old pressure gradient. */
void orc_bcut_remove_p(real* ux, real* uy, real* uz, const real* p, const real* const* c,
                       const int* const* cid, const int* pid, real rhof, real dx, real dt,
                       const int* sz, int g) {
  DECL_SZ
  const ptrdiff_t sy = nxp, sk = nxp * nyp;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        LOAD_GEOM(m)
        face6 r = grad_faces(p, m, 1, sy, sk, 1.0 / rhof, dx, 0, 0.0, 0.0, 0.0, GRAD_ARGS);
        real rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
        ux[m] = ux[m] + rdpx * dt;
        uy[m] = uy[m] + rdpy * dt;
        uz[m] = uz[m] + rdpz * dt;
        if (pidp != 1) { ux[m] = 0.0; uy[m] = 0.0; uz[m] = 0.0; }
      }
}

/* bcut.f90:608-832 bcut_calc_ab_p: Poisson matrix, b = div(v_face)/dt, face velocities. */
void orc_bcut_calc_ab_p(real* const* A, real* b, real* const* v, const real* ux, const real* uy,
                        const real* uz, const real* const* c, const int* const* cid, const int* pid,
                        real rhof, real dx, real dt, const int* sz, int g) {
  DECL_SZ
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real r0 = 1.0 / rhof, r1 = 1.0 / rhof, r2 = 1.0 / rhof, r3 = 1.0 / rhof, r4 = 1.0 / rhof,
             r5 = 1.0 / rhof;
        real vx0 = 0.5 * (ux[m] + ux[IX(i - 1, j, k)]), vx1 = 0.5 * (ux[m] + ux[IX(i + 1, j, k)]);
        real vy2 = 0.5 * (uy[m] + uy[IX(i, j - 1, k)]), vy3 = 0.5 * (uy[m] + uy[IX(i, j + 1, k)]);
        real vz4 = 0.5 * (uz[m] + uz[IX(i, j, k - 1)]), vz5 = 0.5 * (uz[m] + uz[IX(i, j, k + 1)]);
        LOAD_GEOM(m)
        if (cidp0 != 0) { r0 = 0.0; vx0 = vx1 * (d0 - 0.5) / (d0 + 0.5); r1 = (1.0 / rhof) / (d0 + 0.5); }
        if (cidp1 != 0) { r1 = 0.0; vx1 = vx0 * (d1 - 0.5) / (d1 + 0.5); r0 = (1.0 / rhof) / (d1 + 0.5); }
        if (cidp2 != 0) { r2 = 0.0; vy2 = vy3 * (d2 - 0.5) / (d2 + 0.5); r3 = (1.0 / rhof) / (d2 + 0.5); }
        if (cidp3 != 0) { r3 = 0.0; vy3 = vy2 * (d3 - 0.5) / (d3 + 0.5); r2 = (1.0 / rhof) / (d3 + 0.5); }
        if (cidp4 != 0) { r4 = 0.0; vz4 = vz5 * (d4 - 0.5) / (d4 + 0.5); r5 = (1.0 / rhof) / (d4 + 0.5); }
        if (cidp5 != 0) { r5 = 0.0; vz5 = vz4 * (d5 - 0.5) / (d5 + 0.5); r4 = (1.0 / rhof) / (d5 + 0.5); }
        if (cidp0 != 0 && cidp1 != 0) { r0 = 0.0; r1 = 0.0; vx0 = 0.0; vx1 = 0.0; }
        if (cidp2 != 0 && cidp3 != 0) { r2 = 0.0; r3 = 0.0; vy2 = 0.0; vy3 = 0.0; }
        if (cidp4 != 0 && cidp5 != 0) { r4 = 0.0; r5 = 0.0; vz4 = 0.0; vz5 = 0.0; }
        real l0 = r0 / (dx * dx), l1 = r1 / (dx * dx), l2 = r2 / (dx * dx);
        real l3 = r3 / (dx * dx), l4 = r4 / (dx * dx), l5 = r5 / (dx * dx);
        real divv = (vx1 - vx0 + vy3 - vy2 + vz5 - vz4) / dx;
        A[0][m] = -(l0 + l1) - (l2 + l3) - (l4 + l5);
        A[1][m] = l0; A[2][m] = l1; A[3][m] = l2; A[4][m] = l3; A[5][m] = l4; A[6][m] = l5;
        b[m] = divv / dt;
        v[0][m] = vx0; v[1][m] = vx1; v[2][m] = vy2; v[3][m] = vy3; v[4][m] = vz4; v[5][m] = vz5;
        if (pidp != 1) {
          A[0][m] = -6.0;
          for (int q = 1; q < 7; q++) A[q][m] = 1.0;
          b[m] = 0.0;
          for (int q = 0; q < 6; q++) v[q][m] = 0.0;
        }
      }
}

/* bcut.f90:4063-4297 bcut_corr_u: projection, face-velocity correction, lapp (last assignment
 * 4261-4264; zero in solid and cut cells). */
void orc_bcut_corr_u(real* ux, real* uy, real* uz, real* const* v, real* lapp, const real* p,
                     const real* const* c, const int* const* cid, const int* pid, real rhof, real dx,
                     real dt, const int* sz, int g) {
  DECL_SZ
  const ptrdiff_t sy = nxp, sk = nxp * nyp;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        LOAD_GEOM(m)
        face6 r = grad_faces(p, m, 1, sy, sk, 1.0 / rhof, dx, 0, 0.0, 0.0, 0.0, GRAD_ARGS);
        real rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
        ux[m] = ux[m] - rdpx * dt;
        uy[m] = uy[m] - rdpy * dt;
        uz[m] = uz[m] - rdpz * dt;
        v[0][m] = v[0][m] - r.x0 * dt;
        v[1][m] = v[1][m] - r.x1 * dt;
        v[2][m] = v[2][m] - r.y2 * dt;
        v[3][m] = v[3][m] - r.y3 * dt;
        v[4][m] = v[4][m] - r.z4 * dt;
        v[5][m] = v[5][m] - r.z5 * dt;
        lapp[m] = (p[m + 1] + p[m - 1] + p[m + sy] + p[m - sy] + p[m + sk] + p[m - sk] - 6.0 * p[m]) /
                  (dx * dx);
        if (pidp != 1) {
          ux[m] = 0.0; uy[m] = 0.0; uz[m] = 0.0;
          for (int q = 0; q < 6; q++) v[q][m] = 0.0;
          lapp[m] = 0.0;
        }
        if (cidp0 != 0 || cidp1 != 0 || cidp2 != 0 || cidp3 != 0 || cidp4 != 0 || cidp5 != 0) lapp[m] = 0.0;
      }
}

/* bcut.f90:5626-5710 bcut_set_reference_value, including the reference's Ae(i+1,...) index at
 * 5682 (the i-1 neighbour's diagonal is reduced by Ae of the i+1 cell). */
void orc_bcut_set_reference_value(real* const* A, real* b, real xr, real yr, real zr, real pr, real dx,
                                  const real* org, const int* sz, int g) {
  DECL_SZ
  real* Ap = A[0]; real* Aw = A[1]; real* Ae = A[2]; real* As = A[3];
  real* An = A[4]; real* Ab = A[5]; real* At = A[6];
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        real x0 = org[0] + (real)(i - 1) * dx, y0 = org[1] + (real)(j - 1) * dx, z0 = org[2] + (real)(k - 1) * dx;
        real x1 = org[0] + (real)i * dx, y1 = org[1] + (real)j * dx, z1 = org[2] + (real)k * dx;
        if ((x0 <= xr && xr <= x1) && (y0 <= yr && yr <= y1) && (z0 <= zr && zr <= z1)) {
          ptrdiff_t m = IX(i, j, k);
          Ap[m] = 1.0; Aw[m] = 0.0; Ae[m] = 0.0; As[m] = 0.0; An[m] = 0.0; Ab[m] = 0.0; At[m] = 0.0;
          b[m] = pr;
          ptrdiff_t e = IX(i + 1, j, k), w = IX(i - 1, j, k), n = IX(i, j + 1, k), s = IX(i, j - 1, k);
          ptrdiff_t t = IX(i, j, k + 1), bb = IX(i, j, k - 1);
          Ap[e] = Ap[e] - Aw[e]; b[e] = b[e] - Aw[e] * pr * 2.0; Aw[e] = 0.0;
          Ap[w] = Ap[w] - Ae[e]; b[w] = b[w] - Ae[w] * pr * 2.0; Ae[w] = 0.0;
          Ap[n] = Ap[n] - As[n]; b[n] = b[n] - As[n] * pr * 2.0; As[n] = 0.0;
          Ap[s] = Ap[s] - An[s]; b[s] = b[s] - An[s] * pr * 2.0; An[s] = 0.0;
          Ap[t] = Ap[t] - Ab[t]; b[t] = b[t] - Ab[t] * pr * 2.0; Ab[t] = 0.0;
          Ap[bb] = Ap[bb] - At[bb]; b[bb] = b[bb] - At[bb] * pr * 2.0; At[bb] = 0.0;
        }
      }
}

/* bcut.f90:4666-4838 bcut_calc_f_p: pressure force per cut id (the last assignment of each
 * face pressure is the live one, e.g. 4757). */
void orc_bcut_calc_f_p(real* fspx, real* fspy, real* fspz, real* fsp, int cid_target, const real* p,
                       const real* const* c, const int* const* cid, const int* pid, real dx, real dt,
                       const int* sz, int g) {
  DECL_SZ
  (void)dt; (void)c;
  real fsx = 0.0, fsy = 0.0, fsz = 0.0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real pp = p[m];
        real pw = p[IX(i - 1, j, k)], pe = p[IX(i + 1, j, k)], ps = p[IX(i, j - 1, k)];
        real pn = p[IX(i, j + 1, k)], pb = p[IX(i, j, k - 1)], pt = p[IX(i, j, k + 1)];
        real p0 = 0.5 * (pp + pw), p1 = 0.5 * (pp + pe), p2 = 0.5 * (pp + ps);
        real p3 = 0.5 * (pp + pn), p4 = 0.5 * (pp + pb), p5 = 0.5 * (pp + pt);
        real m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, m4 = 0.0, m5 = 0.0;
        if (cid[0][m] == cid_target) { p0 = pp + (pp - pe) * 0.5; m0 = 1.0; }
        if (cid[1][m] == cid_target) { p1 = pp + (pp - pw) * 0.5; m1 = 1.0; }
        if (cid[2][m] == cid_target) { p2 = pp + (pp - pn) * 0.5; m2 = 1.0; }
        if (cid[3][m] == cid_target) { p3 = pp + (pp - ps) * 0.5; m3 = 1.0; }
        if (cid[4][m] == cid_target) { p4 = pp + (pp - pt) * 0.5; m4 = 1.0; }
        if (cid[5][m] == cid_target) { p5 = pp + (pp - pb) * 0.5; m5 = 1.0; }
        if (pid[m] != 1) { m0 = m1 = m2 = m3 = m4 = m5 = 0.0; }
        fspx[m] = (m1 * p1 - m0 * p0);
        fspy[m] = (m3 * p3 - m2 * p2);
        fspz[m] = (m5 * p5 - m4 * p4);
        fsx = fsx + fspx[m] * dx * dx;
        fsy = fsy + fspy[m] * dx * dx;
        fsz = fsz + fspz[m] * dx * dx;
      }
  fsp[0] = fsx; fsp[1] = fsy; fsp[2] = fsz;
}

/* bcut.f90:4840-5032 bcut_calc_f_v: viscous force on the faces cut by boundary id cid_target.
 * A cut face replaces the neighbour velocity by the wall-extrapolated (1 - 1/d) u_p (no-slip wall
 * at distance d); solid cells contribute nothing.  Per component the six face terms are summed
 * left to right as written in the reference; rho, dt and Uc are accepted and unused there. */
void orc_bcut_calc_f_v(real* fsvx, real* fsvy, real* fsvz, real* fsv, int cid_target, const real* ux,
                       const real* uy, const real* uz, const real* const* c, const int* const* cid,
                       const int* pid, real rho, real mu, real dx, real dt, real Uc, const int* sz, int g) {
  DECL_SZ
  (void)rho; (void)dt; (void)Uc;
  real fsx = 0.0, fsy = 0.0, fsz = 0.0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        const ptrdiff_t nb[6] = {IX(i - 1, j, k), IX(i + 1, j, k), IX(i, j - 1, k),
                                 IX(i, j + 1, k), IX(i, j, k - 1), IX(i, j, k + 1)};
        real up[3] = {ux[m], uy[m], uz[m]};
        real un[3][6];  /* component, face (w e s n b t) */
        const real* u3[3] = {ux, uy, uz};
        for (int a = 0; a < 3; a++)
          for (int f = 0; f < 6; f++) un[a][f] = u3[a][nb[f]];
        real mf[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < 6; f++)
          if (cid[f][m] == cid_target) {
            const real w = 1.0 - 1.0 / c[f][m];
            for (int a = 0; a < 3; a++) un[a][f] = w * up[a];
            mf[f] = 1.0;
          }
        if (pid[m] != 1)
          for (int f = 0; f < 6; f++) mf[f] = 0.0;
        /* component a: its own axis uses the doubled normal gradient; the two other axes add half
         * the cross difference of the tangential component along a */
        real* out[3] = {fsvx, fsvy, fsvz};
        for (int a = 0; a < 3; a++) {
          real acc = 0.0;
          for (int ax = 0; ax < 3; ax++) {
            const int fm = 2 * ax, fp = 2 * ax + 1;
            real lo, hi;
            if (ax == a) {
              lo = mu * ((up[a] - un[a][fm])) / dx * mf[fm] * 2.0;
              hi = mu * ((un[a][fp] - up[a])) / dx * mf[fp] * 2.0;
            } else {
              /* d u_ax / d x_a from the two neighbours along a */
              const real cross = (un[ax][2 * a + 1] - un[ax][2 * a]) * 0.5;
              lo = mu * ((up[a] - un[a][fm]) + cross) / dx * mf[fm];
              hi = mu * ((un[a][fp] - up[a]) + cross) / dx * mf[fp];
            }
            acc = (ax == 0) ? +lo - hi : acc + lo - hi;
          }
          out[a][m] = acc;
        }
        fsx = fsx + fsvx[m] * dx * dx;
        fsy = fsy + fsvy[m] * dx * dx;
        fsz = fsz + fsvz[m] * dx * dx;
      }
  fsv[0] = fsx; fsv[1] = fsy; fsv[2] = fsz;
}

/* ------------------------------------------------------------------ bstl.f90 */

/* bstl.f90:1306-1372 bstl_cutoff */
void orc_bstl_cutoff(real* const* c, int* const* cid, real eps, const int* sz, int g) {
  DECL_SZ
  (void)cid;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        for (int d = 0; d < 6; d++) c[d][m] = fmax(eps, c[d][m]);
      }
}

/* bstl.f90:1557-1771 bstl_detect_zerocut */
int orc_bstl_detect_zerocut(real* const* c, int* const* cid, const int* sz, int g) {
  DECL_SZ
  int n = 0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t m = IX(i, j, k);
        real cut_min = 1.0;
        int cid_min = 0;
        for (int d = 0; d < 6; d++)
          if (cut_min > c[d][m]) { cut_min = c[d][m]; cid_min = cid[d][m]; }
        if (cut_min == 0.0) {
          for (int d = 0; d < 6; d++) c[d][m] = 1.0;
          c[1][IX(i - 1, j, k)] = 1.0; cid[1][IX(i - 1, j, k)] = cid_min;
          c[0][IX(i + 1, j, k)] = 1.0; cid[0][IX(i + 1, j, k)] = cid_min;
          c[3][IX(i, j - 1, k)] = 1.0; cid[3][IX(i, j - 1, k)] = cid_min;
          c[2][IX(i, j + 1, k)] = 1.0; cid[2][IX(i, j + 1, k)] = cid_min;
          c[5][IX(i, j, k - 1)] = 1.0; cid[5][IX(i, j, k - 1)] = cid_min;
          c[4][IX(i, j, k + 1)] = 1.0; cid[4][IX(i, j, k + 1)] = cid_min;
          n = n + 1;
        }
      }
  return n;
}

/* bstl.f90:1773-2023 bstl_fill_holes_v1 (all four side tests must hold) and 2025-2275
 * bstl_fill_holes_v2 ((first two) or (last two)): close a cut face whose side faces are cut.
 * Single sweep in (k,j,i) order; later tests in a cell see the faces closed earlier in it. */
static int fill_holes(real* const* c, int* const* cid, int bClose, int v2, const int* sz, int g) {
  DECL_SZ
  int n = 0;
  for (int k = 1; k <= kx; k++)
    for (int j = 1; j <= jx; j++)
      for (int i = 1; i <= ix; i++) {
        ptrdiff_t p = IX(i, j, k), w = IX(i - 1, j, k), e = IX(i + 1, j, k), s = IX(i, j - 1, k);
        ptrdiff_t nn = IX(i, j + 1, k), b = IX(i, j, k - 1), t = IX(i, j, k + 1);
#define C(d, q) (cid[d][q] != 0)
#define HOLE(a, bb, cc, dd) (v2 ? (((a) && (bb)) || ((cc) && (dd))) : ((a) && (bb) && (cc) && (dd)))
        if (cid[0][p] == 0) {
          if (HOLE((C(0, s) || C(2, p) || C(2, w)), (C(0, nn) || C(3, p) || C(3, w)),
                   (C(0, b) || C(4, p) || C(4, w)), (C(0, t) || C(5, p) || C(5, w)))) {
            if (bClose == 1) { cid[0][p] = 1; cid[1][w] = 1; c[0][p] = 0.5; c[1][w] = 0.5; }
            n++;
          }
        }
        if (cid[1][p] == 0) {
          if (HOLE((C(1, s) || C(2, p) || C(2, e)), (C(1, nn) || C(3, p) || C(3, e)),
                   (C(1, b) || C(4, p) || C(4, e)), (C(1, t) || C(5, p) || C(5, e)))) {
            if (bClose == 1) { cid[1][p] = 1; cid[0][e] = 1; c[1][p] = 0.5; c[0][e] = 0.5; }
            n++;
          }
        }
        if (cid[2][p] == 0) {
          if (HOLE((C(2, w) || C(0, p) || C(0, s)), (C(2, e) || C(1, p) || C(1, s)),
                   (C(2, b) || C(4, p) || C(4, s)), (C(2, t) || C(5, p) || C(5, s)))) {
            if (bClose == 1) { cid[2][p] = 1; cid[3][s] = 1; c[2][p] = 0.5; c[3][s] = 0.5; }
            n++;
          }
        }
        if (cid[3][p] == 0) {
          if (HOLE((C(3, w) || C(0, p) || C(0, nn)), (C(3, e) || C(1, p) || C(1, nn)),
                   (C(3, b) || C(4, p) || C(4, nn)), (C(3, t) || C(5, p) || C(5, nn)))) {
            if (bClose == 1) { cid[3][p] = 1; cid[2][nn] = 1; c[3][p] = 0.5; c[2][nn] = 0.5; }
            n++;
          }
        }
        if (cid[4][p] == 0) {
          if (HOLE((C(4, w) || C(0, p) || C(0, b)), (C(4, e) || C(1, p) || C(1, b)),
                   (C(4, s) || C(2, p) || C(2, b)), (C(4, nn) || C(3, p) || C(3, b)))) {
            if (bClose == 1) { cid[4][p] = 1; cid[5][b] = 1; c[4][p] = 0.5; c[5][b] = 0.5; }
            n++;
          }
        }
        if (cid[5][p] == 0) {
          if (HOLE((C(5, w) || C(0, p) || C(0, t)), (C(5, e) || C(1, p) || C(1, t)),
                   (C(5, s) || C(2, p) || C(2, t)), (C(5, nn) || C(3, p) || C(3, t)))) {
            if (bClose == 1) { cid[5][p] = 1; cid[4][t] = 1; c[5][p] = 0.5; c[4][t] = 0.5; }
            n++;
          }
        }
#undef HOLE
#undef C
      }
  return n;
}

int orc_bstl_fill_holes_v1(real* const* c, int* const* cid, int bClose, const int* sz, int g) {
  return fill_holes(c, cid, bClose, 0, sz, g);
}

int orc_bstl_fill_holes_v2(real* const* c, int* const* cid, int bClose, const int* sz, int g) {
  return fill_holes(c, cid, bClose, 1, sz, g);
}
