// Device helpers shared by the solver kernels: deterministic per-case reductions, the
// convergence test and the halo push epilogue.
#pragma once

#include <cfloat>

#include "device.h"
#include "ghost.h"

namespace ffvhc {

// ------------------------------------------------------------------------------------------
// Per-case deterministic reductions: every workgroup writes its block's partial sum(s); the
// per-case finalisation adds the partials in slot order with a fixed tree.
// ------------------------------------------------------------------------------------------

template <int NT>
__device__ inline double block_sum(double v, double* red) {
  // fixed-pattern wave reduction, then a fixed-order pass over the waves
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < NT / 64; q++) s += red[q];
  return s;  // valid in thread 0
}

template <int NT>
__device__ inline double case_sum(const double* __restrict__ part, int begin, int count, double* red) {
  // strided per-thread sums, then the fixed wave pattern and the waves in order (deterministic)
  double v = 0.0;
#pragma unroll 4
  for (int q = threadIdx.x; q < count; q += NT) v += part[begin + q];  // loads issued 4 at a time
  return block_sum<NT>(v, red);  // valid in thread 0
}

// two sums over the same partial range in one pass (both valid in thread 0)
template <int NT>
__device__ inline void case_sum2(const double* __restrict__ p0, const double* __restrict__ p1, int begin, int count,
                                 double* red, double& s0, double& s1) {
  double v0 = 0.0, v1 = 0.0;
#pragma unroll 4
  for (int q = threadIdx.x; q < count; q += NT) {
    v0 += p0[begin + q];
    v1 += p1[begin + q];
  }
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
  }
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[w] = v0;
    red[NT / 64 + w] = v1;
  }
  __syncthreads();
  s0 = 0.0;
  s1 = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < NT / 64; q++) {
      s0 += red[q];
      s1 += red[NT / 64 + q];
    }
}

// Planes a thread loads per batch: every global load of a batch is issued before any store of
// the batch, so the memory latency is paid once per batch instead of once per plane (a store
// through a pointer held in a struct may alias later loads as far as the compiler can tell).
template <class T>
struct Batch {
  static constexpr int K = sizeof(T) == 4 ? 8 : 4;
};

constexpr int LP = NB + 2;  // LDS pitch
__device__ inline int lds(int i, int j, int k) { return (i + 1) + LP * ((j + 1) + LP * (k + 1)); }

// Stage a block's interior and its one-cell ghost layer into LDS (pitch LP); NT threads with
// thread t owning column (t & 15, t >> 4) of every plane (NT = 256).
template <class T, int NT>
__device__ inline void stage_block(T* s, const T* __restrict__ x, const T* __restrict__ h, int slot) {
  static_assert(NT == 256, "one thread per (i, j) column");
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  T v[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) v[k] = x[base + tx + NB * (ty + NB * k)];
  T hv[NH / NT];
#pragma unroll
  for (int q = 0; q < NH / NT; q++) hv[q] = h[(size_t)slot * NH + t + q * NT];
#pragma unroll
  for (int k = 0; k < NB; k++) s[lds(tx, ty, k)] = v[k];
#pragma unroll
  for (int q = 0; q < NH / NT; q++) {
    const int w = t + q * NT;
    int f = w >> 8, p = w & 255, i, j, kk;
    face_cell(f, p & 15, p >> 4, -1, i, j, kk);
    s[lds(i, j, kk)] = hv[q];
  }
}


__device__ inline void converged(SolveState& s, double rr, int count) {
  // FFVILS.h:24-47 IsConverged
  s.rr = rr;
  double bb = s.bb;
  double residual = (fabs(bb) > FLT_MIN) ? sqrt(fabs(rr) / fabs(bb)) : sqrt(fabs(rr));
  s.residual = residual;
  s.count = count;
  if (residual <= s.eps) {
    s.done = 1;
  } else if (isnan(residual) || isinf(residual)) {
    s.done = 1;
    s.status = 1;
  } else if (count >= s.countMax) {
    s.done = 1;
    s.count = s.countMax + 1;  // loop exhausted: for(count=1; count<=countMax; ++count)
  }
}

// case_sum2<256> formed by one wavefront: lane l stands for the threads l, l + 64, l + 128 and
// l + 192 of the 256-thread form (the same strided sums in the same order, the same butterfly per
// wave, the four waves added in order), so the result has the same bits; valid in every lane.
__device__ inline void case_sum2_wave(const double* __restrict__ p0, const double* __restrict__ p1, int begin,
                                      int count, double& s0, double& s1) {
  const int lane = threadIdx.x & 63;
  double v0[4] = {0.0, 0.0, 0.0, 0.0}, v1[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int q0 = 0; q0 < count; q0 += 256) {  // loads issued four rounds at a time
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const int q = q0 + lane + 64 * w;
      if (q < count) {
        v0[w] += p0[begin + q];
        v1[w] += p1[begin + q];
      }
    }
  }
  s0 = 0.0;
  s1 = 0.0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    for (int off = 32; off > 0; off >>= 1) {
      v0[w] += __shfl_xor(v0[w], off, 64);
      v1[w] += __shfl_xor(v1[w], off, 64);
    }
    s0 += v0[w];
    s1 += v1[w];
  }
}

// The end of a BiCGSTAB iteration (FFVILS.h:925-965): IsConverged on t0 = r.r, then the next
// rho t1 = r.r0 and beta = (t1 / rho_n) (alpha / gamma), rho_n being this iteration's rho.  The
// evaluation reads only fields of S that no kernel writes while it runs, so any workgroup can form
// it; fin_bi_xr_store writes it as the per-case finalisation does.
struct FinXr {
  double rr, residual, rr1, beta;
  int count, done, status;
};

__device__ inline FinXr fin_bi_xr_eval(const SolveState& S, double t0, double t1, int n, double rho_n) {
  FinXr f;
  f.rr = t0;
  const double bb = S.bb;
  f.residual = (fabs(bb) > FLT_MIN) ? sqrt(fabs(t0) / fabs(bb)) : sqrt(fabs(t0));
  f.count = n;
  f.done = 0;
  f.status = 0;
  if (f.residual <= S.eps) {
    f.done = 1;
  } else if (isnan(f.residual) || isinf(f.residual)) {
    f.done = 1;
    f.status = 1;
  } else if (n >= S.countMax) {
    f.done = 1;
    f.count = S.countMax + 1;
  }
  f.rr1 = 0.0;
  f.beta = 0.0;
  if (!f.done) {
    if (fabs(t1) < FLT_MIN) {  // next iteration's r.r0 test (FFVILS.h:829-834)
      f.residual = t1;
      f.count = 0;
      f.done = 1;
    } else {
      f.rr1 = t1;
      f.beta = t1 / rho_n * S.alpha / S.gamma;
    }
  }
  return f;
}

__device__ inline void fin_bi_xr_store(SolveState& S, const FinXr& f, int n, double rho_n) {
  S.iter = n;
  S.rr = f.rr;
  S.residual = f.residual;
  S.count = f.count;
  if (f.status) S.status = 1;
  S.rr0 = rho_n;  // this iteration's rho becomes the previous one (also kept for a reliable update)
  if (f.done) {
    S.done = 1;
  } else {
    S.rr1 = f.rr1;
    S.beta = f.beta;
  }
}


// Neighbour slot a face position (a, b) of face f takes its ghost value from.
__device__ inline int ghost_source(const DevBlock& B, int slot, int f, int a, int b) {
  switch (B.ftype[f]) {
    case FACE_SAME:
    case FACE_COARSER:
      return B.nbr[f][0];
    case FACE_FINER: {
      int u0, u1;
      face_ab_to_o(f, a, b, u0, u1);
      return B.nbr[f][(u0 >= NB / 2 ? 1 : 0) + (u1 >= NB / 2 ? 2 : 0)];
    }
    default:
      return slot;  // outer face: the Neumann ghost of the block's own cell
  }
}

// Push this block's ghost-layer contributions into the next halo buffer: its own outer faces
// (bcx.f90), the facing layer of same-level neighbours, F2C means into a coarser neighbour's
// quarter face and C2F trilinear values into each finer neighbour (Scalar3DUpdater1 semantics).
// `L(i,j,k)` reads the block's new interior (LDS).
template <class T, int NT, class Acc>
__device__ inline void push_halo(const DevBlock& B, const DevCase& C, int bcg, int slot, const Acc& L,
                                 const T* __restrict__ hin, T* __restrict__ hout) {
  const int t = threadIdx.x;
  const size_t hbase = (size_t)slot * NH;
  for (int f = 0; f < 6; f++) {
    const int ft = B.ftype[f];
    if (ft == FACE_OUTER) {
      const int type = C.bct[bcg][f];
      if (type < 0) continue;
      T xc = (T)(type == BCT_NEUMANN ? C.bcv[bcg][f] * B.dx : C.bcv[bcg][f]);
      bool pp = !(f & 1) && B.ftype[f ^ 1] == FACE_OUTER && C.bct[bcg][f ^ 1] == BCT_PERIODIC;
      for (int p = t; p < NF; p += NT) {
        int a = p & 15, b = p >> 4, i, j, k;
        face_cell(f, a, b, 0, i, j, k);
        T xa = L(i, j, k);
        face_cell(f ^ 1, a, b, 0, i, j, k);
        T xf = L(i, j, k);
        hout[hbase + f * NF + p] = outer_value<T>(f, type, xc, xa, xf, hin[hbase + f * NF + p], pp);
      }
    } else if (ft == FACE_SAME) {
      const size_t nb = (size_t)B.nbr[f][0] * NH + (f ^ 1) * NF;
      for (int p = t; p < NF; p += NT) {
        int i, j, k;
        face_cell(f, p & 15, p >> 4, 0, i, j, k);
        hout[nb + p] = L(i, j, k);
      }
    } else if (ft == FACE_COARSER) {
      const size_t cb = (size_t)B.nbr[f][0] * NH;
      int oa, ob;
      sub_origins_ab(f, B.sub[f], oa, ob);
      for (int p = t; p < NF / 4; p += NT) {
        int ca = (p & 7) + 8 * oa, cbb = (p >> 3) + 8 * ob;
        hout[cb + halo_index(f ^ 1, ca, cbb)] = f2c_mean<T>(f ^ 1, ca, cbb, L, false);
      }
    } else {
      for (int q = t; q < 4 * NF; q += NT) {
        const int sub = q >> 8, p = q & 255;
        const size_t fb = (size_t)B.nbr[f][sub] * NH;
        hout[fb + halo_index(f ^ 1, p & 15, p >> 4)] = c2f_trilinear<T>(f ^ 1, sub, p & 15, p >> 4, L);
      }
    }
  }
}

}  // namespace ffvhc
