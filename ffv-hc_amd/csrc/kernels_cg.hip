// Jacobi-preconditioned CG (FFVILS::CG, FFVILS.h:661-791) as three fused kernels per iteration,
// with every scalar (alpha, beta, the convergence test) computed on the device by the per-case
// finalisation (kernels_reduce.hip).  The preconditioner is the reference's Jacobi_PreConditioner with
// countPreConditioner = 1 from z = 0 (FFVILS.h:468-490, bils.f90:5-55):
// z = 0 - omega*(0 - (r - A_nb*0 - ...)/Ap), which equals omega*(r/Ap) exactly in IEEE arithmetic
// (every A_nb*0 is a signed zero, x - (+-0) = x, 0 - (0 - y) = y), so only Ap is read.
//
//   init : r = b - A x (calc_r), b.b, z, r.z                                      [FFVILS.h:679-712]
//   P    : p = z (first iteration, Copy) or z + beta p (XPAY); push p's ghost layer [716-731]
//   Q    : q = A p (calc_ax), q.p -> alpha                                        [733-749]
//   XR   : x += alpha p, r -= alpha q, r.r -> IsConverged; next z, r.z -> beta     [751-787]
#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

constexpr int CG_NT = 256;

template <class T>
__device__ inline T jacobi_z(T r, T ap, T omega) {
  return omega * (r / ap);
}

template <class T>
__global__ __launch_bounds__(CG_NT) void cg_init_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                        const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                        Coef7<T> A, const T* __restrict__ bv, const T* __restrict__ x,
                                                        const T* __restrict__ hx, T* __restrict__ r,
                                                        double* __restrict__ part, size_t nslots) {
  __shared__ T s[LP * LP * LP];
  __shared__ double red[CG_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const T omega = (T)st[c].omega;
  constexpr int KB = Batch<T>::K;
  double bb = 0.0, rz = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    T a[KB][7], b0[KB];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
#pragma unroll
      for (int q = 0; q < 7; q++) a[kk][q] = A.p[q][g];
      b0[kk] = bv[g];
    }
    if (k0 == 0) {
      stage_block<T, CG_NT>(s, x, hx, slot);
      __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      const int li = lds(tx, ty, k0 + kk);
      const T rr = b0[kk] - a[kk][1] * s[li - 1] - a[kk][2] * s[li + 1] - a[kk][3] * s[li - LP] -
                   a[kk][4] * s[li + LP] - a[kk][5] * s[li - LP * LP] - a[kk][6] * s[li + LP * LP] -
                   a[kk][0] * s[li];
      r[g] = rr;
      const T z = jacobi_z<T>(rr, a[kk][0], omega);
      bb += (double)b0[kk] * (double)b0[kk];
      rz += (double)rr * (double)z;
    }
  }
  double sb = block_sum<CG_NT>(bb, red);
  double sz = block_sum<CG_NT>(rz, red);
  if (t == 0) {
    part[slot] = sb;
    part[nslots + slot] = sz;
  }
}

template <class T>
__global__ __launch_bounds__(CG_NT) void cg_p_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                     const DevCase* __restrict__ cs, const SolveState* __restrict__ st,
                                                     Coef7<T> A, const T* __restrict__ r, T* __restrict__ p,
                                                     T* __restrict__ hp, int count, int bcg) {
  __shared__ T s[LP * LP * LP];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const T omega = (T)st[c].omega;
  const T beta = (T)st[c].beta;
  constexpr int KB = Batch<T>::K;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    T rv[KB], ap[KB], pv[KB];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      rv[kk] = r[g];
      ap[kk] = A.p[0][g];
      if (count != 1) pv[kk] = p[g];
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      const T z = jacobi_z<T>(rv[kk], ap[kk], omega);
      const T pn = (count == 1) ? z : z + beta * pv[kk];
      p[g] = pn;
      s[lds(tx, ty, k0 + kk)] = pn;
    }
  }
  __syncthreads();
  push_halo<T, CG_NT>(B, cs[c], bcg, slot, [&](int i, int j, int k) { return s[lds(i, j, k)]; }, hp, hp);
}

template <class T>
__global__ __launch_bounds__(CG_NT) void cg_q_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                     const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                     Coef7<T> A, const T* __restrict__ p, const T* __restrict__ hp,
                                                     T* __restrict__ q, double* __restrict__ part) {
  __shared__ T s[LP * LP * LP];
  __shared__ double red[CG_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  constexpr int KB = Batch<T>::K;
  double qp = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    T a[KB][7];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
#pragma unroll
      for (int w = 0; w < 7; w++) a[kk][w] = A.p[w][g];
    }
    if (k0 == 0) {
      stage_block<T, CG_NT>(s, p, hp, slot);
      __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      const int li = lds(tx, ty, k0 + kk);
      const T pv = s[li];
      const T qv = a[kk][1] * s[li - 1] + a[kk][2] * s[li + 1] + a[kk][3] * s[li - LP] + a[kk][4] * s[li + LP] +
                   a[kk][5] * s[li - LP * LP] + a[kk][6] * s[li + LP * LP] + a[kk][0] * pv;
      q[g] = qv;
      qp += (double)qv * (double)pv;
    }
  }
  double sq = block_sum<CG_NT>(qp, red);
  if (t == 0) part[slot] = sq;
}

template <class T>
__global__ __launch_bounds__(CG_NT) void cg_xr_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                      const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                      Coef7<T> A, T* __restrict__ x, T* __restrict__ r,
                                                      const T* __restrict__ p, const T* __restrict__ q,
                                                      double* __restrict__ part, size_t nslots, int count) {
  __shared__ double red[CG_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const T alpha = (T)st[c].alpha;
  const T nalpha = -alpha;
  const T omega = (T)st[c].omega;
  constexpr int KB = Batch<T>::K;
  double rr = 0.0, rz = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    T xv[KB], pv[KB], rv[KB], qv[KB], ap[KB];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      xv[kk] = x[g];
      pv[kk] = p[g];
      rv[kk] = r[g];
      qv[kk] = q[g];
      ap[kk] = A.p[0][g];
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      x[g] = xv[kk] + alpha * pv[kk];
      const T rn = rv[kk] + nalpha * qv[kk];
      r[g] = rn;
      const T z = jacobi_z<T>(rn, ap[kk], omega);
      rr += (double)rn * (double)rn;
      rz += (double)rn * (double)z;
    }
  }
  double sr = block_sum<CG_NT>(rr, red);
  double sz = block_sum<CG_NT>(rz, red);
  if (t == 0) {
    part[slot] = sr;
    part[nslots + slot] = sz;
  }
}

template <class T>
void launch_cg_init(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs, SolveState* st,
                    Coef7<T> A, const T* b, const T* x, const T* hx, T* r, double* part, size_t nslots) {
  if (nact) hipLaunchKernelGGL(cg_init_kernel<T>, dim3(nact), dim3(CG_NT), 0, sm, active, blk, cs, st, A, b, x, hx, r,
                               part, nslots);
}
template <class T>
void launch_cg_p(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                 const SolveState* st, Coef7<T> A, const T* r, T* p, T* hp, int count, int bcg) {
  if (nact) hipLaunchKernelGGL(cg_p_kernel<T>, dim3(nact), dim3(CG_NT), 0, sm, active, blk, cs, st, A, r, p, hp,
                               count, bcg);
}
template <class T>
void launch_cg_q(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs, SolveState* st,
                 Coef7<T> A, const T* p, const T* hp, T* q, double* part) {
  if (nact) hipLaunchKernelGGL(cg_q_kernel<T>, dim3(nact), dim3(CG_NT), 0, sm, active, blk, cs, st, A, p, hp, q, part);
}
template <class T>
void launch_cg_xr(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs, SolveState* st,
                  Coef7<T> A, T* x, T* r, const T* p, const T* q, double* part, size_t nslots, int count) {
  if (nact) hipLaunchKernelGGL(cg_xr_kernel<T>, dim3(nact), dim3(CG_NT), 0, sm, active, blk, cs, st, A, x, r, p, q,
                               part, nslots, count);
}

#define INST(T)                                                                                                   \
  template void launch_cg_init<T>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, SolveState*,      \
                                  Coef7<T>, const T*, const T*, const T*, T*, double*, size_t);                   \
  template void launch_cg_p<T>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, const SolveState*,   \
                               Coef7<T>, const T*, T*, T*, int, int);                                              \
  template void launch_cg_q<T>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, SolveState*,         \
                               Coef7<T>, const T*, const T*, T*, double*);                                         \
  template void launch_cg_xr<T>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, SolveState*,        \
                                Coef7<T>, T*, T*, const T*, const T*, double*, size_t, int);
INST(double)
INST(float)
#undef INST

}  // namespace ffvhc

namespace ffvhc {

// ------------------------------------------------------------------------------------------
// FP64 defect correction around an FP32 inner solve.
//   defect: r = b - A x in FP64 (calc_r order, x with its ghost layer), r.r and (first pass) b.b,
//           IsConverged on the FP64 residual; r is handed to the inner solve as FP32.
//   update: x += e (FP32 correction widened to FP64) over the interior.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(CG_NT) void defect_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                       const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                       Coef7<double> A, const double* __restrict__ bv,
                                                       const double* __restrict__ x, const double* __restrict__ hx,
                                                       float* __restrict__ r32, double* __restrict__ part,
                                                       size_t nslots, int outer) {
  __shared__ double s[LP * LP * LP];
  __shared__ double red[CG_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  constexpr int KB = Batch<double>::K;
  double rr = 0.0, bb = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    double a[KB][7], b0[KB];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
#pragma unroll
      for (int q = 0; q < 7; q++) a[kk][q] = A.p[q][g];
      b0[kk] = bv[g];
    }
    if (k0 == 0) {
      stage_block<double, CG_NT>(s, x, hx, slot);
      __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      const int li = lds(tx, ty, k0 + kk);
      const double r = b0[kk] - a[kk][1] * s[li - 1] - a[kk][2] * s[li + 1] - a[kk][3] * s[li - LP] -
                       a[kk][4] * s[li + LP] - a[kk][5] * s[li - LP * LP] - a[kk][6] * s[li + LP * LP] -
                       a[kk][0] * s[li];
      r32[g] = (float)r;
      rr += r * r;
      if (outer == 0) bb += b0[kk] * b0[kk];
    }
  }
  double sr = block_sum<CG_NT>(rr, red);
  double sb = outer == 0 ? block_sum<CG_NT>(bb, red) : 0.0;
  if (t == 0) {
    part[slot] = sr;
    part[nslots + slot] = sb;
  }
}

__global__ __launch_bounds__(256) void add_correction_kernel(const int* __restrict__ active,
                                                             const DevBlock* __restrict__ blk,
                                                             SolveState* __restrict__ st,
                                                             const SolveState* __restrict__ inner,
                                                             double* __restrict__ x, const float* __restrict__ e) {
  const int slot = active[blockIdx.x];
  const int c = blk[slot].case_id;
  if (st[c].done) return;
  if (inner[c].status) {  // the inner solve broke down (NaN/Inf): the case ends with that status
    if (threadIdx.x == 0) {
      st[c].status = 1;
      st[c].residual = inner[c].residual;
      st[c].done = 1;
    }
    return;
  }
  const size_t base = (size_t)slot * NC;
  for (int q = threadIdx.x; q < NC; q += 256) x[base + q] = x[base + q] + (double)e[base + q];
}

__global__ __launch_bounds__(256) void convert_kernel(const int* __restrict__ active, float* __restrict__ y,
                                                      const double* __restrict__ x) {
  const int slot = active[blockIdx.x];
  const size_t base = (size_t)slot * NC;
  for (int q = threadIdx.x; q < NC; q += 256) y[base + q] = (float)x[base + q];
}

void launch_defect(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                   SolveState* st, Coef7<double> A, const double* b, const double* x, const double* hx, float* r32,
                   double* part, size_t nslots, int outer) {
  if (nact) hipLaunchKernelGGL(defect_kernel, dim3(nact), dim3(CG_NT), 0, sm, active, blk, cs, st, A, b, x, hx, r32,
                               part, nslots, outer);
}
void launch_add_correction(hipStream_t sm, const int* active, int nact, const DevBlock* blk, SolveState* st,
                           const SolveState* inner, double* x, const float* e) {
  if (nact) hipLaunchKernelGGL(add_correction_kernel, dim3(nact), dim3(256), 0, sm, active, blk, st, inner, x, e);
}
void launch_convert(hipStream_t sm, const int* active, int nact, float* y, const double* x) {
  if (nact) hipLaunchKernelGGL(convert_kernel, dim3(nact), dim3(256), 0, sm, active, y, x);
}

}  // namespace ffvhc
