/*
 * ffv_oracle_grid.c -- grid-level part of the CPU oracle (test infrastructure only):
 * ghost-layer fill over a block table, the FFVILS linear solvers and Solver::Update.
 *
 * Per-block loops run OpenMP-parallel over blocks, as the reference does with
 * FFV_THREAD=BLOCK (src/make_setting:43-44); global sums add per-block partial
 * sums in block order, so the oracle is deterministic (the reference's
 * `omp reduction` order is not).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ffv_oracle.h"

#define NPAD(gr) ((size_t)((gr)->n + 2 * (gr)->g) * ((gr)->n + 2 * (gr)->g) * ((gr)->n + 2 * (gr)->g))
#define BLK(f, gr, b) ((f) + (size_t)(b) * NPAD(gr))

/* Planted mistakes (test infrastructure): the pin tests switch one on to show that their check
 * fails when the oracle is wrong.  0 = none (the only value outside those tests). */
static int plant_id = 0;
void orc_plant(int id) { plant_id = id; }
#define PLANTED(id) (plant_id == (id))

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ ghosts */

/* Scalar3DUpdater1.cpp:432-523 (F2C), 11-99 (same level), 247-350 (C2F) index setup. */
static void u1_f2c(const orc_grid* gr, real* dst, const real* src, int face, int s) {
  const int n = gr->n, vc = gr->g, sz[3] = {n, n, n};
  const int o0 = s & 1, o1 = s >> 1;
  int d[3] = {1, 1, 1}, r[3] = {1, 1, 1}, c[3] = {n / 2, n / 2, n / 2};
  switch (face) {
    case 0: d[0] = 1 - vc; d[1] = 1 + (n / 2) * o0; d[2] = 1 + (n / 2) * o1; r[0] = 1 + n / 2 - vc; c[0] = vc; break;
    case 1: d[0] = 1 + n;  d[1] = 1 + (n / 2) * o0; d[2] = 1 + (n / 2) * o1; c[0] = vc; break;
    case 2: d[0] = 1 + (n / 2) * o1; d[1] = 1 - vc; d[2] = 1 + (n / 2) * o0; r[1] = 1 + n / 2 - vc; c[1] = vc; break;
    case 3: d[0] = 1 + (n / 2) * o1; d[1] = 1 + n;  d[2] = 1 + (n / 2) * o0; c[1] = vc; break;
    case 4: d[0] = 1 + (n / 2) * o0; d[1] = 1 + (n / 2) * o1; d[2] = 1 - vc; r[2] = 1 + n / 2 - vc; c[2] = vc; break;
    case 5: d[0] = 1 + (n / 2) * o0; d[1] = 1 + (n / 2) * o1; d[2] = 1 + n;  c[2] = vc; break;
  }
  orc_sup_copy_from_neighbor_f2c(dst, d, src, r, c, sz, vc);
}

static void u1_same(const orc_grid* gr, real* dst, const real* src, int face) {
  const int n = gr->n, vc = gr->g, sz[3] = {n, n, n};
  int d[3] = {1, 1, 1}, r[3] = {1, 1, 1}, c[3] = {n, n, n};
  const int axis = face / 2;
  c[axis] = vc;
  if (face & 1) { d[axis] = 1 + n; r[axis] = 1; }
  else { d[axis] = 1 - vc; r[axis] = 1 + n - vc; }
  orc_sup_copy_from_neighbor(dst, d, src, r, c, sz, vc);
}

static void u1_c2f(const orc_grid* gr, real* dst, const real* src, int face, int s) {
  const int n = gr->n, vc = gr->g, sz[3] = {n, n, n};
  const int o0 = s & 1, o1 = s >> 1;
  int d[3] = {1, 1, 1}, r[3] = {1, 1, 1}, c[3] = {n, n, n};
  switch (face) {
    case 0: d[0] = 1 - vc; r[0] = 1 + 2 * n - vc; r[1] = 1 + n * o0; r[2] = 1 + n * o1; c[0] = vc; break;
    case 1: d[0] = 1 + n;  r[0] = 1;              r[1] = 1 + n * o0; r[2] = 1 + n * o1; c[0] = vc; break;
    case 2: d[1] = 1 - vc; r[0] = 1 + n * o1; r[1] = 1 + 2 * n - vc; r[2] = 1 + n * o0; c[1] = vc; break;
    case 3: d[1] = 1 + n;  r[0] = 1 + n * o1; r[1] = 1;              r[2] = 1 + n * o0; c[1] = vc; break;
    case 4: d[2] = 1 - vc; r[0] = 1 + n * o0; r[1] = 1 + n * o1; r[2] = 1 + 2 * n - vc; c[2] = vc; break;
    case 5: d[2] = 1 + n;  r[0] = 1 + n * o0; r[1] = 1 + n * o1; r[2] = 1;              c[2] = vc; break;
  }
  orc_sup_copy_from_neighbor_c2f(dst, d, src, r, c, sz, vc);
}

/* LocalScalar3D.h:266-296 ImposeBoundaryCondition: virtual-cell update then the outer-face BC.
 * The VC update is done in three passes, F2C -> same level -> C2F (the order made explicit by
 * SeparateXYZ_SeparateLevelDiff, LocalScalar3D.h:252-262); Scalar3DUpdater5's C2F reads the
 * coarse ghost that F2C filled (Scalar3DUpdater5.h:785-786), so F2C must come first. */
static void ghost_c2f(const orc_grid* gr, real* f, int updater) {
  const int sz[3] = {gr->n, gr->n, gr->n};
#pragma omp parallel for schedule(static)
  for (int b = 0; b < gr->nblk; b++)
    for (int face = 0; face < 6; face++)
      if (gr->ftype[b * 6 + face] == 2) {
        int src = gr->nbr[(b * 6 + face) * 4], s = gr->sub[b * 6 + face];
        if (updater == 5) orc_u5_c2f(BLK(f, gr, b), BLK(f, gr, src), face, s, sz, gr->g);
        else u1_c2f(gr, BLK(f, gr, b), BLK(f, gr, src), face, s);
      }
}

void orc_ghost_fill(const orc_grid* gr, real* f, int updater, const int* bctype, const real* bcval) {
  const int nb = gr->nblk;
  const int sz[3] = {gr->n, gr->n, gr->n};
  if (PLANTED(8)) ghost_c2f(gr, f, updater);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; b++)
    for (int face = 0; face < 6; face++)
      if (gr->ftype[b * 6 + face] == 3)
        for (int s = 0; s < 4; s++) {
          int src = gr->nbr[(b * 6 + face) * 4 + s];
          if (updater == 5) orc_u5_f2c(BLK(f, gr, b), BLK(f, gr, src), face, s, sz, gr->g);
          else u1_f2c(gr, BLK(f, gr, b), BLK(f, gr, src), face, s);
        }
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; b++)
    for (int face = 0; face < 6; face++)
      if (gr->ftype[b * 6 + face] == 1) u1_same(gr, BLK(f, gr, b), BLK(f, gr, gr->nbr[(b * 6 + face) * 4]), face);
  if (!PLANTED(8)) ghost_c2f(gr, f, updater);
  /* BlockScalar3D.h:174-181 in face order X_M X_P Y_M Y_P Z_M Z_P; Neumann value * dx (74-76) */
  if (!bctype) return;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; b++)
    for (int face = 0; face < 6; face++)
      if (gr->ftype[b * 6 + face] == 0 && bctype[face] >= 0) {
        real xc = bcval[face];
        if (bctype[face] == 1) xc = bcval[face] * gr->dx[b];
        orc_bc_x(BLK(f, gr, b), face, bctype[face], xc, sz, gr->g);
      }
}

/* LocalScalar3D.h:298-323 -> BlockScalar3D.h:183-190: Aw Ae As An Ab At in turn. */
void orc_matrix_bc(const orc_grid* gr, real* const* A, real* b, const int* bctype, const real* bcval) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  for (int blk = 0; blk < gr->nblk; blk++)
    for (int face = 0; face < 6; face++)
      if (gr->ftype[blk * 6 + face] == 0 && (bctype[face] == 0 || bctype[face] == 1)) {
        real xc = bcval[face];
        if (bctype[face] == 1) xc = bcval[face] * gr->dx[blk];
        orc_bc_a(BLK(A[0], gr, blk), BLK(A[1 + face], gr, blk), BLK(b, gr, blk), face, bctype[face], xc,
                 sz, gr->g);
      }
}

/* ------------------------------------------------------------ reductions */

static real sum_partials(const real* part, int nb) {
  real s = 0.0;
  for (int b = 0; b < nb; b++) s += part[b];
  return s;
}

/* FFVILS.h:298-328 DOT */
real orc_grid_dot(const orc_grid* gr, const real* x, const real* y) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  real* part = (real*)malloc(sizeof(real) * (size_t)gr->nblk);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < gr->nblk; b++) part[b] = orc_dot(BLK(x, gr, b), BLK(y, gr, b), sz, gr->g);
  real s = sum_partials(part, gr->nblk);
  free(part);
  return s;
}

/* FFVILS.h:247-296 CalcR2 */
real orc_grid_r2(const orc_grid* gr, real* const* A, const real* x, const real* b) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  real* part = (real*)malloc(sizeof(real) * (size_t)gr->nblk);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < gr->nblk; k++)
    part[k] = orc_calc_r2(BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k), BLK(A[3], gr, k),
                          BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k), BLK(x, gr, k),
                          BLK(b, gr, k), sz, gr->g);
  real s = sum_partials(part, gr->nblk);
  free(part);
  return s;
}

/* LocalScalar3D.h:130-159 CalcSum (serial block loop) */
real orc_grid_sum(const orc_grid* gr, const real* x) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  real s = 0.0;
  for (int b = 0; b < gr->nblk; b++) s += orc_sf3d_calc_sum(BLK(x, gr, b), sz, gr->g);
  return s;
}

/* FFVILS.h:24-47 IsConverged; returns 1 converged, 0 not, -1 NaN/Inf (reference Exit(0)) */
static int is_converged(real* residual, real rr, real bb, real eps) {
  *residual = (fabs(bb) > FLT_MIN) ? sqrt(fabs(rr) / fabs(bb)) : sqrt(fabs(rr));
  if (*residual <= eps) return 1;
  if (isnan(*residual) || isinf(*residual)) return -1;
  return 0;
}

/* --------------------------------------------------------------- solvers */

/* FFVILS.h:104-162 RBGS_Smoother */
static void rbgs_sweep(const orc_grid* gr, real* x, real* const* A, const real* b, real omega,
                       const int* bctype, const real* bcval) {
  const int sz[3] = {gr->n, gr->n, gr->n};
#pragma omp parallel for schedule(static)
  for (int k = 0; k < gr->nblk; k++)
    for (int color = 0; color < 2; color++)
      orc_rbgs_smoother(BLK(x, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k),
                        BLK(A[3], gr, k), BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k),
                        BLK(b, gr, k), omega, color, 0, sz, gr->g);
  orc_ghost_fill(gr, x, 1, bctype, bcval);
}

/* FFVILS.h:547-594 RBGS */
int orc_rbgs(const orc_grid* gr, real* x, real* const* A, const real* b, const int* bctype,
             const real* bcval, real omega, int countMax, real eps, int* count, real* residual) {
  real bb = orc_grid_dot(gr, b, b);
  int c;
  for (c = 1; c <= countMax; ++c) {
    rbgs_sweep(gr, x, A, b, omega, bctype, bcval);
    real rr = orc_grid_r2(gr, A, x, b);
    int st = is_converged(residual, rr, bb, eps);
    if (st < 0) { *count = c; return 1; }
    if (st > 0) break;
  }
  *count = c;
  return 0;
}

/* FFVILS.h:596-659 RBGS_WithFixedAverage (CalcSum / ShiftByA, LocalScalar3D.h:130-182) */
int orc_rbgs_fixed_average(const orc_grid* gr, real* x, real* const* A, const real* b,
                           const int* bctype, const real* bcval, int Ncells, real omega, int countMax,
                           real eps, int* count, real* residual) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  real bb = orc_grid_dot(gr, b, b);
  int c;
  for (c = 1; c <= countMax; ++c) {
    rbgs_sweep(gr, x, A, b, omega, bctype, bcval);
    real sum_x1 = orc_grid_sum(gr, x);
    real a = -sum_x1 / (double)Ncells;
    if (PLANTED(4)) a = 0.0;
    for (int k = 0; k < gr->nblk; k++) orc_adda(BLK(x, gr, k), a, sz, gr->g);
    orc_ghost_fill(gr, x, 1, bctype, bcval);
    real rr = orc_grid_r2(gr, A, x, b);
    int st = is_converged(residual, rr, bb, eps);
    if (st < 0) { *count = c; return 1; }
    if (st > 0) break;
  }
  *count = c;
  return 0;
}

/* FFVILS.h:49-102 Jacobi_Smoother: x0 <- jacobi(x); swap; ghost fill.  (The reference's
 * outer-face BC after the swap lands in the pre-swap array, Appendix-A quirk of LSSwap; the
 * oracle fills the live array.  Outer links are zero after the matrix BC, so no value read by
 * the solver differs.) */
static void jacobi_sweep(const orc_grid* gr, real** x, real* const* A, const real* b, real** x0,
                         real omega, const int* bctype, const real* bcval) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  real* xs = *x;
  real* x0s = *x0;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < gr->nblk; k++)
    orc_jacobi_smoother(BLK(x0s, gr, k), BLK(xs, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k),
                        BLK(A[2], gr, k), BLK(A[3], gr, k), BLK(A[4], gr, k), BLK(A[5], gr, k),
                        BLK(A[6], gr, k), BLK(b, gr, k), omega, sz, gr->g);
  if (!PLANTED(3)) { /* planted: no swap, the sweep's result is dropped */
    *x = x0s;
    *x0 = xs;
  }
  orc_ghost_fill(gr, *x, 1, bctype, bcval);
}

/* FFVILS.h:492-545 Jacobi.  On return x holds the solution (copied back if the swaps left it in
 * the scratch array). */
int orc_jacobi(const orc_grid* gr, real* x, real* const* A, const real* b, real* x0, const int* bctype,
               const real* bcval, real omega, int countMax, real eps, int* count, real* residual) {
  real bb = orc_grid_dot(gr, b, b);
  real* xa = x;
  real* xb = x0;
  int c, ret = 0;
  for (c = 1; c <= countMax; ++c) {
    jacobi_sweep(gr, &xa, A, b, &xb, omega, bctype, bcval);
    real rr = orc_grid_r2(gr, A, xa, b);
    int st = is_converged(residual, rr, bb, eps);
    if (st < 0) { ret = 1; break; }
    if (st > 0) break;
  }
  orc_ghost_fill(gr, xa, 1, bctype, bcval);
  if (xa != x) memcpy(x, xa, sizeof(real) * NPAD(gr) * (size_t)gr->nblk);
  *count = c;
  return ret;
}

/* FFVILS.h:661-791 CG (Jacobi-preconditioned; z from countPre damped-Jacobi sweeps, z = 0 start) */
int orc_cg(const orc_grid* gr, real* x, real* const* A, const real* b, real* r, real* p, real* q, real* z,
           real* x0, const int* bctype, const real* bcval, real omega, int countPre, int countMax,
           real eps, int* count, real* residual) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  const int nb = gr->nblk;
  /* scratch fields carry Neumann-0 outer BC (boundaryTypeNULL, Solver.cpp:2024-2034) */
  const int bt_null[6] = {1, 1, 1, 1, 1, 1};
  const real bv_null[6] = {0, 0, 0, 0, 0, 0};
  (void)bctype;
  real bb = orc_grid_dot(gr, b, b);
  if (fabs(bb) < FLT_MIN) { *residual = 0.0; *count = 0; return 0; }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; k++)
    orc_calc_r(BLK(r, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k), BLK(A[3], gr, k),
               BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k), BLK(x, gr, k), BLK(b, gr, k), sz, gr->g);
  orc_ghost_fill(gr, r, 1, bt_null, bv_null);
  real rr0 = 1.0, rr1 = 0.0;
  real* zz = z;
  real* xx0 = x0;
  int c, ret = 0;
  for (c = 1; c <= countMax; ++c) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) orc_fill(BLK(zz, gr, k), 0.0, sz, gr->g);
    for (int it = 0; it < countPre; it++) jacobi_sweep(gr, &zz, A, r, &xx0, omega, bt_null, bv_null);
    rr1 = orc_grid_dot(gr, r, zz);
    if (fabs(rr1) < FLT_MIN) { *residual = rr1; c = 0; break; }
    real beta = PLANTED(7) ? rr0 / rr1 : rr1 / rr0;
    if (c == 1) {
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nb; k++) orc_copy(BLK(p, gr, k), BLK(zz, gr, k), sz, gr->g);
    } else {
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nb; k++) orc_xpay(BLK(p, gr, k), BLK(zz, gr, k), beta, sz, gr->g);
    }
    orc_ghost_fill(gr, p, 1, bt_null, bv_null);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++)
      orc_calc_ax(BLK(q, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k), BLK(A[3], gr, k),
                  BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k), BLK(p, gr, k), sz, gr->g);
    real qp = orc_grid_dot(gr, q, p);
    real alpha = rr1 / qp;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) {
      orc_axpy(BLK(x, gr, k), BLK(p, gr, k), alpha, sz, gr->g);
      orc_axpy(BLK(r, gr, k), BLK(q, gr, k), -alpha, sz, gr->g);
    }
    orc_ghost_fill(gr, r, 1, bt_null, bv_null);
    real rr = orc_grid_dot(gr, r, r);
    int st = is_converged(residual, rr, bb, eps);
    if (st < 0) { ret = 1; break; }
    if (st > 0) break;
    rr0 = rr1;
  }
  orc_ghost_fill(gr, x, 1, bctype, bcval);
  *count = c;
  return ret;
}

/* FFVILS::BiCGSTAB (FFVILS.h:793-974): Jacobi right-preconditioned BiCGSTAB; scratch r, r0, p, p_,
 * q_, s, s_, t_ (Solver.cpp:2026-2042) and x0 (the preconditioner's swap array). */
int orc_bicgstab(const orc_grid* gr, real* x, real* const* A, const real* b, real* const* sc, const int* bctype,
                 const real* bcval, real omega, int countPre, int countMax, real eps, int* count, real* residual) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  const int nb = gr->nblk;
  const int bt_null[6] = {1, 1, 1, 1, 1, 1};
  const real bv_null[6] = {0, 0, 0, 0, 0, 0};
  real *r = sc[0], *r0 = sc[1], *p = sc[2], *q_ = sc[4], *s = sc[5], *t_ = sc[7];
  real *p_ = sc[3], *s_ = sc[6], *x0 = sc[8];
  real bb = orc_grid_dot(gr, b, b);
  if (fabs(bb) < FLT_MIN) { *residual = 0.0; *count = 0; return 0; }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; k++)
    orc_calc_r(BLK(r, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k), BLK(A[3], gr, k),
               BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k), BLK(x, gr, k), BLK(b, gr, k), sz, gr->g);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; k++) orc_copy(BLK(r0, gr, k), BLK(r, gr, k), sz, gr->g);
  real rr0 = 1.0, alpha = 0.0, gamma = 1.0;
  int c, ret = 0;
  for (c = 1; c <= countMax; ++c) {
    real rr1 = orc_grid_dot(gr, r, r0);
    if (fabs(rr1) < FLT_MIN) { *residual = rr1; c = 0; break; }
    if (c == 1) {
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nb; k++) orc_copy(BLK(p, gr, k), BLK(r, gr, k), sz, gr->g);
    } else {
      real beta = rr1 / rr0 * alpha / gamma;
      if (PLANTED(1)) beta = -beta;
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nb; k++) {
        orc_axpy(BLK(p, gr, k), BLK(q_, gr, k), -gamma, sz, gr->g);
        orc_xpay(BLK(p, gr, k), BLK(r, gr, k), beta, sz, gr->g);
      }
    }
    orc_ghost_fill(gr, p, 1, bt_null, bv_null);
    real* pp = p_;
    real* xx0 = x0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) orc_fill(BLK(pp, gr, k), 0.0, sz, gr->g);
    for (int it = 0; it < countPre; it++) jacobi_sweep(gr, &pp, A, p, &xx0, omega, bt_null, bv_null);
    if (pp != p_) { /* keep p_ naming stable for the caller-visible scratch */
      memcpy(p_, pp, sizeof(real) * NPAD(gr) * (size_t)nb);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++)
      orc_calc_ax(BLK(q_, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k), BLK(A[3], gr, k),
                  BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k), BLK(p_, gr, k), sz, gr->g);
    real q_r0 = orc_grid_dot(gr, q_, r0);
    alpha = rr1 / q_r0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) orc_axpyz(BLK(s, gr, k), BLK(q_, gr, k), BLK(r, gr, k), -alpha, sz, gr->g);
    orc_ghost_fill(gr, s, 1, bt_null, bv_null);
    real* ss = s_;
    xx0 = x0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) orc_fill(BLK(ss, gr, k), 0.0, sz, gr->g);
    for (int it = 0; it < countPre; it++) jacobi_sweep(gr, &ss, A, s, &xx0, omega, bt_null, bv_null);
    if (ss != s_) memcpy(s_, ss, sizeof(real) * NPAD(gr) * (size_t)nb);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++)
      orc_calc_ax(BLK(t_, gr, k), BLK(A[0], gr, k), BLK(A[1], gr, k), BLK(A[2], gr, k), BLK(A[3], gr, k),
                  BLK(A[4], gr, k), BLK(A[5], gr, k), BLK(A[6], gr, k), BLK(s_, gr, k), sz, gr->g);
    real t_s = orc_grid_dot(gr, t_, s);
    real t_t_ = orc_grid_dot(gr, t_, t_);
    gamma = t_s / t_t_;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) {
      if (PLANTED(2)) orc_axpbypz(BLK(x, gr, k), BLK(p_, gr, k), BLK(s_, gr, k), gamma, alpha, sz, gr->g);
      else orc_axpbypz(BLK(x, gr, k), BLK(p_, gr, k), BLK(s_, gr, k), alpha, gamma, sz, gr->g);
      orc_axpyz(BLK(r, gr, k), BLK(t_, gr, k), BLK(s, gr, k), -gamma, sz, gr->g);
    }
    real rr = orc_grid_dot(gr, r, r);
    int st = is_converged(residual, rr, bb, eps);
    if (st < 0) { ret = 1; break; }
    if (st > 0) break;
    rr0 = rr1;
  }
  orc_ghost_fill(gr, x, 1, bctype, bcval);
  *count = c;
  return ret;
}

/* ------------------------------------------------------------ time step */

#define GEOM(F, gr, k)                                                                               \
  const real* cb[6] = {BLK(F->c[0], gr, k), BLK(F->c[1], gr, k), BLK(F->c[2], gr, k),              \
                       BLK(F->c[3], gr, k), BLK(F->c[4], gr, k), BLK(F->c[5], gr, k)};             \
  const int* ib[6] = {F->cid[0] + (size_t)k * NPAD(gr), F->cid[1] + (size_t)k * NPAD(gr),           \
                      F->cid[2] + (size_t)k * NPAD(gr), F->cid[3] + (size_t)k * NPAD(gr),           \
                      F->cid[4] + (size_t)k * NPAD(gr), F->cid[5] + (size_t)k * NPAD(gr)};          \
  const int* pb = F->pid + (size_t)k * NPAD(gr);

/* UpdateU (Solver.cpp:5690-5771): bcut_corr_u per block, then the velocity ghosts. */
static void corr_stage(const orc_grid* gr, orc_fields* F, const orc_params* P, real* const* u,
                       real (*bcv_u)[6]) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  const int nb = gr->nblk, g = gr->g;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; k++) {
    GEOM(F, gr, k)
    real* vb[6];
    for (int q = 0; q < 6; q++) vb[q] = BLK(F->v[q], gr, k);
    orc_bcut_corr_u(BLK(F->ux, gr, k), BLK(F->uy, gr, k), BLK(F->uz, gr, k), vb, BLK(F->lapp, gr, k),
                    BLK(F->p, gr, k), cb, ib, pb, P->rho, gr->dx[k], P->dt, sz, g);
  }
  for (int d = 0; d < 3; d++) orc_ghost_fill(gr, u[d], 5, P->bct_u[d], bcv_u[d]);
}

/* Solver::Update (Solver.cpp:3344-3416) with UpdateUXe/UYe/UZe (3535-3744, explicit U),
 * UpdateP (5294-5688) and UpdateU (5690-5771). */
void orc_update(const orc_grid* gr, orc_fields* F, const orc_params* P, int step, orc_step_info* info) {
  const int sz[3] = {gr->n, gr->n, gr->n};
  const int nb = gr->nblk, g = gr->g;
  const real dt = P->dt, rhof = P->rho, mu = P->mu, Uc = 0.0;
  const int bt_null[6] = {1, 1, 1, 1, 1, 1};
  const real bv_null[6] = {0, 0, 0, 0, 0, 0};
  real bcv_u[3][6];
  memcpy(bcv_u, P->bcv_u, sizeof(bcv_u));
  /* Solver.cpp:3347-3355 inflow ramp on X_M of UX */
  if (P->accel_steps > 0 && step <= P->accel_steps)
    bcv_u[0][0] = P->bcv_u[0][0] * (real)step / (real)P->accel_steps;
  real* u[3] = {F->ux, F->uy, F->uz};
  double t0 = now_s(), tt[7];
  tt[0] = t0;
  tt[1] = now_s(); /* UpdateT: out of scope */

  if (P->u_implicit) {
    /* UpdateUX / UY / UZ (Solver.cpp:4391-4690): calc_c_f_c2, calc_d_u, calc_abd_u, copy uc to
     * ucp per block; matrix BC of the component; BiCGSTAB; ghost fill (Updater5). */
    for (int d = 0; d < 3; d++) {
#pragma omp parallel for schedule(static)
      for (int k = 0; k < nb; k++) {
        GEOM(F, gr, k)
        real dx = gr->dx[k];
        real* uk = BLK(u[d], gr, k);
        real* Ab[7];
        for (int q = 0; q < 7; q++) Ab[q] = BLK(F->A[q], gr, k);
        orc_bcut_calc_c_f_c2(BLK(F->uc[d], gr, k), uk, BLK(F->v[0], gr, k), BLK(F->v[1], gr, k),
                             BLK(F->v[2], gr, k), BLK(F->v[3], gr, k), BLK(F->v[4], gr, k),
                             BLK(F->v[5], gr, k), cb, ib, pb, dx, dt, Uc, sz, g);
        orc_bcut_calc_d_u(BLK(F->ud[d], gr, k), uk, cb, ib, pb, rhof, mu, dx, dt, Uc, sz, g);
        orc_bcut_calc_abd_u(Ab, BLK(F->b, gr, k), uk, BLK(F->uc[d], gr, k), BLK(F->ucp[d], gr, k),
                            BLK(F->ud[d], gr, k), BLK(F->p, gr, k), cb, ib, pb, d, P->alpha_u, rhof, mu, dx, dt,
                            Uc, P->gx, P->gy, P->gz, sz, g);
        orc_copy(BLK(F->ucp[d], gr, k), BLK(F->uc[d], gr, k), sz, g);
      }
      orc_matrix_bc(gr, F->A, F->b, P->bct_u[d], bcv_u[d]);
      int st = orc_bicgstab(gr, u[d], F->A, F->b, F->bis, P->bct_u[d], bcv_u[d], P->omega_u, 1, P->countMax_u,
                            P->eps_u, &info->countU[d], &info->residualU[d]);
      (void)st;
      orc_ghost_fill(gr, u[d], 5, P->bct_u[d], bcv_u[d]);
      tt[2 + d] = now_s();
    }
  } else {
  /* UpdateUXe / UYe / UZe */
  for (int d = 0; d < 3; d++) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; k++) {
      GEOM(F, gr, k)
      real dx = gr->dx[k];
      real* uk = BLK(u[d], gr, k);
      orc_bcut_calc_c_f_c2(BLK(F->uc[d], gr, k), uk, BLK(F->v[0], gr, k), BLK(F->v[1], gr, k),
                           BLK(F->v[2], gr, k), BLK(F->v[3], gr, k), BLK(F->v[4], gr, k),
                           BLK(F->v[5], gr, k), cb, ib, pb, dx, dt, Uc, sz, g);
      orc_bcut_calc_d_u(BLK(F->ud[d], gr, k), uk, cb, ib, pb, rhof, mu, dx, dt, Uc, sz, g);
      if (step == 0) {
        orc_copy(BLK(F->ucp[d], gr, k), BLK(F->uc[d], gr, k), sz, g);
        orc_copy(BLK(F->udp[d], gr, k), BLK(F->ud[d], gr, k), sz, g);
      }
      orc_bcut_update_u(uk, BLK(F->uc[d], gr, k), BLK(F->ucp[d], gr, k), BLK(F->ud[d], gr, k),
                        BLK(F->udp[d], gr, k), BLK(F->p, gr, k), cb, ib, pb, d, rhof, mu, dx, dt, Uc,
                        P->gx, P->gy, P->gz, sz, g);
      orc_copy(BLK(F->ucp[d], gr, k), BLK(F->uc[d], gr, k), sz, g);
      orc_copy(BLK(F->udp[d], gr, k), BLK(F->ud[d], gr, k), sz, g);
    }
    orc_ghost_fill(gr, u[d], 5, P->bct_u[d], bcv_u[d]);
    tt[2 + d] = now_s();
  }

  }

  /* UpdateP */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; k++) {
    GEOM(F, gr, k)
    real dx = gr->dx[k];
    real *ux = BLK(F->ux, gr, k), *uy = BLK(F->uy, gr, k), *uz = BLK(F->uz, gr, k);
    orc_bcut_add_g(ux, uy, uz, BLK(F->t0, gr, k), P->gx, P->gy, P->gz, P->betag, P->tr, dx, dt, sz, g);
    orc_bcut_add_f(ux, uy, uz, BLK(F->f[0], gr, k), BLK(F->f[1], gr, k), BLK(F->f[2], gr, k), pb, rhof,
                   dx, dt, sz, g);
    if (!PLANTED(6)) orc_bcut_remove_p(ux, uy, uz, BLK(F->p, gr, k), cb, ib, pb, rhof, dx, dt, sz, g);
  }
  for (int d = 0; d < 3; d++) orc_ghost_fill(gr, u[d], 5, P->bct_u[d], bcv_u[d]);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nb; k++) {
    GEOM(F, gr, k)
    real* Ab[7];
    real* vb[6];
    for (int q = 0; q < 7; q++) Ab[q] = BLK(F->A[q], gr, k);
    for (int q = 0; q < 6; q++) vb[q] = BLK(F->v[q], gr, k);
    orc_bcut_calc_ab_p(Ab, BLK(F->b, gr, k), vb, BLK(F->ux, gr, k), BLK(F->uy, gr, k), BLK(F->uz, gr, k),
                       cb, ib, pb, rhof, gr->dx[k], dt, sz, g);
  }
  if (P->ref_active) {
    for (int k = 0; k < nb; k++) {
      real* Ab[7];
      for (int q = 0; q < 7; q++) Ab[q] = BLK(F->A[q], gr, k);
      orc_bcut_set_reference_value(Ab, BLK(F->b, gr, k), P->ref_x, P->ref_y, P->ref_z, P->ref_p, gr->dx[k],
                                   gr->org + 3 * k, sz, g);
    }
  }
  orc_matrix_bc(gr, F->A, F->b, P->bct_p, P->bcv_p);
  if (PLANTED(5)) corr_stage(gr, F, P, u, bcv_u);
  int st = 0;
  if (P->solver == 0)
    st = orc_rbgs(gr, F->p, F->A, F->b, P->bct_p, P->bcv_p, P->omega, P->countMax, P->eps, &info->countP,
                  &info->residualP);
  else if (P->solver == 1)
    st = orc_rbgs_fixed_average(gr, F->p, F->A, F->b, P->bct_p, P->bcv_p, P->ncells, P->omega, P->countMax,
                                P->eps, &info->countP, &info->residualP);
  else if (P->solver == 2)
    st = orc_cg(gr, F->p, F->A, F->b, F->cgs[0], F->cgs[1], F->cgs[2], F->cgs[3], F->cgs[4], P->bct_p,
                P->bcv_p, P->omega, P->countPre, P->countMax, P->eps, &info->countP, &info->residualP);
  else /* Jacobi-BiCGSTAB (SolvingMethod/P = BiCGSTAB, Solver.cpp:5582-5677) */
    st = orc_bicgstab(gr, F->p, F->A, F->b, F->bis, P->bct_p, P->bcv_p, P->omega, P->countPre, P->countMax,
                      P->eps, &info->countP, &info->residualP);
  orc_ghost_fill(gr, F->p, 1, P->bct_p, P->bcv_p);
  tt[5] = now_s();

  /* UpdateU */
  if (!PLANTED(5)) corr_stage(gr, F, P, u, bcv_u);
  orc_ghost_fill(gr, F->p, 1, P->bct_p, P->bcv_p);
  orc_ghost_fill(gr, F->lapp, 1, bt_null, bv_null);
  tt[6] = now_s();

  info->status = st;
  info->times[0] = tt[6] - tt[0];
  info->times[1] = tt[1] - tt[0];
  for (int q = 2; q <= 6; q++) info->times[q] = tt[q] - tt[q - 1];
}
