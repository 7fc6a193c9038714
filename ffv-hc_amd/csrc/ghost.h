// Ghost-layer values: outer-face BC, same-level copy, fine-to-coarse restriction and
// coarse-to-fine prolongation, written once and used both by the pull kernel (every field)
// and by the push epilogue of the fused solver kernels.
//
//   same level  : Scalar3DUpdater1.cpp:11-99 / sup.f90:18-52
//   F2C mean    : sup.f90:122-179 (sum order nnn pnn npn nnp npp pnp ppn ppp) for updater 1,
//                 Scalar3DUpdater5.h:198-203 (nnn pnn npn ppn nnp pnp npp ppp) for updater 5
//   C2F, upd. 1 : sup.f90:1-16 + 54-120, trilinear from the coarse interior, extrapolated at ends
//   C2F, upd. 5 : Scalar3DUpdater5.h:765-885, f_ghost = f_adjacent + 0.5*(c_in - c_ghost)
//   outer BC    : bcx.f90 (D 5-147, N 294-436, P 585-658), face map BlockScalar3D.h:201-254
#pragma once

#include "device.h"

namespace ffvhc {

// axes spanned by a face's (a,b) coordinates
__host__ __device__ inline void face_axes(int face, int& ax_a, int& ax_b) {
  switch (face >> 1) {
    case 0: ax_a = 1; ax_b = 2; break;
    case 1: ax_a = 0; ax_b = 2; break;
    default: ax_a = 0; ax_b = 1; break;
  }
}

// subface origins (o0,o1) re-expressed along the face's (a,b) axes
__host__ __device__ inline void sub_origins_ab(int face, int sub, int& oa, int& ob) {
  const int o0 = sub & 1, o1 = sub >> 1;
  if ((face >> 1) == 1) { oa = o1; ob = o0; }  // Y faces: a = x (origin1), b = z (origin0)
  else { oa = o0; ob = o1; }
}

// sup_get_intp_index on 1-based doubled index i (0..2n+1), returns 0-based coarse index
__host__ __device__ inline void intp_index(int i, int& ic0, double& r) {
  int ic = i / 2;
  r = 0.25 + 0.5 * (double)(i % 2);
  if (i < 2) { ic = 1; r = -0.25; }
  else if (i > 2 * NB - 1) { ic = NB - 1; r = 1.25; }
  ic0 = ic - 1;
}

// Mean of the 8 fine cells under coarse-face position (a,b) of coarse face `fc`.
// `fine(i,j,k)` reads the fine block's interior.
template <class T, class Src>
__device__ inline T f2c_mean(int fc, int a, int b, const Src& fine, bool u5) {
  int ax_a, ax_b;
  face_axes(fc, ax_a, ax_b);
  int lo[3];
  lo[fc >> 1] = (fc & 1) ? 0 : NB - 2;
  lo[ax_a] = 2 * (a % (NB / 2));
  lo[ax_b] = 2 * (b % (NB / 2));
  const int i = lo[0], j = lo[1], k = lo[2];
  T s;
  if (!u5) {
    s = fine(i, j, k) + fine(i + 1, j, k) + fine(i, j + 1, k) + fine(i, j, k + 1) + fine(i, j + 1, k + 1) +
        fine(i + 1, j, k + 1) + fine(i + 1, j + 1, k) + fine(i + 1, j + 1, k + 1);
  } else {
    s = fine(i, j, k) + fine(i + 1, j, k) + fine(i, j + 1, k) + fine(i + 1, j + 1, k) + fine(i, j, k + 1) +
        fine(i + 1, j, k + 1) + fine(i, j + 1, k + 1) + fine(i + 1, j + 1, k + 1);
  }
  return (T)0.125 * s;
}

// Trilinear C2F value for fine face `ff` (coarse neighbour across it, this block is quarter
// `sub` of the coarse face) at fine-face position (a,b); `coarse(i,j,k)` reads the coarse interior.
template <class T, class Src>
__device__ inline T c2f_trilinear(int ff, int sub, int a, int b, const Src& coarse) {
  int ax_a, ax_b, oa, ob;
  face_axes(ff, ax_a, ax_b);
  sub_origins_ab(ff, sub, oa, ob);
  int ic[3];
  double w[3];
  const int nrm = ff >> 1;
  if (ff & 1) { ic[nrm] = 0; w[nrm] = -0.25; }
  else { ic[nrm] = NB - 2; w[nrm] = 1.25; }
  intp_index(a + 1 + NB * oa, ic[ax_a], w[ax_a]);
  intp_index(b + 1 + NB * ob, ic[ax_b], w[ax_b]);
  const int i = ic[0], j = ic[1], k = ic[2];
  const T r = (T)w[0], s = (T)w[1], t = (T)w[2];
  const T one = (T)1.0;
  T nnn = coarse(i, j, k), pnn = coarse(i + 1, j, k), npn = coarse(i, j + 1, k), nnp = coarse(i, j, k + 1);
  T npp = coarse(i, j + 1, k + 1), pnp = coarse(i + 1, j, k + 1), ppn = coarse(i + 1, j + 1, k);
  T ppp = coarse(i + 1, j + 1, k + 1);
  T a0 = (one - r) * nnn + r * pnn;
  T a1 = (one - r) * npn + r * ppn;
  T a2 = (one - r) * nnp + r * pnp;
  T a3 = (one - r) * npp + r * ppp;
  T c0 = (one - s) * a0 + s * a1;
  T c1 = (one - s) * a2 + s * a3;
  return (one - t) * c0 + t * c1;
}

// coarse-face position that lies across fine-face position (a,b)
__host__ __device__ inline void c2f_coarse_ab(int ff, int sub, int a, int b, int& ca, int& cb) {
  int oa, ob;
  sub_origins_ab(ff, sub, oa, ob);
  ca = a / 2 + (NB / 2) * oa;
  cb = b / 2 + (NB / 2) * ob;
}

// Outer-face ghost (layer 1).  xc is already scaled (Neumann value * dx).
template <class T>
__device__ inline T outer_value(int face, int type, T xc, T x_adj, T x_far, T x_old, bool plus_periodic_on_minus) {
  const bool plus = face & 1;
  if (!plus && plus_periodic_on_minus) return x_far;  // bc_x1_p writes x(0) = x(n)
  switch (type) {
    case BCT_DIRICHLET: return (T)2.0 * xc - x_adj;
    case BCT_NEUMANN: return plus ? x_adj + xc : x_adj - xc;
    case BCT_PERIODIC: return plus ? x_far : x_old;  // minus-side periodic is a no-op
    default: return x_old;
  }
}

}  // namespace ffvhc
