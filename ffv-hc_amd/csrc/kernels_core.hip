// Core kernels: ghost-layer pull fill, fused RBGS iteration (residual + red/black sweep +
// halo push + per-case convergence), result selection, BLAS-1 helpers.
#include <cfloat>

#include "device.h"
#include "ghost.h"
#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

// ------------------------------------------------------------------------------------------
// Pull ghost fill (LocalScalar3D::ImposeBoundaryCondition, LocalScalar3D.h:266-296).
// phase 0: outer BC, same level, F2C;  phase 1: C2F (updater 5 reads the coarse ghost that
// phase 0 wrote).  One workgroup per (block, face), one thread per face cell.
// ------------------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void ghost_pull_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                          const DevCase* __restrict__ cs, const T* __restrict__ x,
                                                          T* __restrict__ h, int bcg, int updater, int phase) {
  const int slot = active[blockIdx.x / 6];
  const int f = blockIdx.x % 6;
  const DevBlock& B = blk[slot];
  const int ft = B.ftype[f];
  if ((phase == 0) == (ft == FACE_COARSER)) return;
  const int p = threadIdx.x, a = p & (NB - 1), b = p >> 4;
  const T* me = x + (size_t)slot * NC;
  auto own = [&](int i, int j, int k) { return me[i + NB * (j + NB * k)]; };
  int i, j, k;
  T* hp = h + (size_t)slot * NH + halo_index(f, a, b);
  T v;
  if (ft == FACE_OUTER) {
    const DevCase& C = cs[B.case_id];
    const int type = C.bct[bcg][f];
    if (type < 0) return;
    T xc = (T)(type == BCT_NEUMANN ? C.bcv[bcg][f] * B.dx : C.bcv[bcg][f]);
    face_cell(f, a, b, 0, i, j, k);
    T x_adj = own(i, j, k);
    face_cell(f ^ 1, a, b, 0, i, j, k);
    T x_far = own(i, j, k);
    bool pp = !(f & 1) && B.ftype[f ^ 1] == FACE_OUTER && C.bct[bcg][f ^ 1] == BCT_PERIODIC;
    v = outer_value<T>(f, type, xc, x_adj, x_far, *hp, pp);
  } else if (ft == FACE_SAME) {
    const T* nb = x + (size_t)B.nbr[f][0] * NC;
    face_cell(f ^ 1, a, b, 0, i, j, k);
    v = nb[i + NB * (j + NB * k)];
  } else if (ft == FACE_FINER) {
    int oa = a / (NB / 2), ob = b / (NB / 2);
    int s = ((f >> 1) == 1) ? (ob + 2 * oa) : (oa + 2 * ob);
    const T* fb = x + (size_t)B.nbr[f][s] * NC;
    auto fine = [&](int ii, int jj, int kk) { return fb[ii + NB * (jj + NB * kk)]; };
    v = f2c_mean<T>(f, a, b, fine, updater == 5);
  } else {  // FACE_COARSER
    const int sub = B.sub[f];
    const int cslot = B.nbr[f][0];
    const T* cb = x + (size_t)cslot * NC;
    auto coarse = [&](int ii, int jj, int kk) { return cb[ii + NB * (jj + NB * kk)]; };
    if (updater == 5) {
      int ca, cbb;
      c2f_coarse_ab(f, sub, a, b, ca, cbb);
      face_cell(f ^ 1, ca, cbb, 0, i, j, k);
      T c_in = coarse(i, j, k);
      T c_gh = h[(size_t)cslot * NH + halo_index(f ^ 1, ca, cbb)];
      T dp = (c_in - c_gh) * (T)0.5;
      face_cell(f, a, b, 0, i, j, k);
      v = own(i, j, k) + dp;
    } else {
      v = c2f_trilinear<T>(f, sub, a, b, coarse);
    }
  }
  *hp = v;
}

// ------------------------------------------------------------------------------------------
// Fused RBGS iteration n (FFVILS::RBGS, FFVILS.h:547-594 with RBGS_Smoother 104-162 and
// rbgs_smoother bils.f90:57-108):
//   * r = b - A x^{n-1} over the block (CalcR2 of the previous iterate; iteration 1 instead
//     accumulates b.b, the DOT before the loop),
//   * red then black SOR on block-local parity with the ghost layer frozen,
//   * x^n to the other ping-pong buffer, its ghost layer pushed into the neighbours' next halo
//     (same-level copy, F2C mean, C2F trilinear: Scalar3DUpdater1 semantics) and the outer BC.
// The per-case finalisation completes the residual of x^{n-1} and sets `done`; later
// launches of a finished case return at once, so the device never waits for the host.
// ------------------------------------------------------------------------------------------
constexpr int RB_NT = 512;

// MODE 0: FFVILS::RBGS iteration.  MODE 1: FFVILS::Jacobi iteration (jacobi_smoother,
// bils.f90:5-55: every cell from the old values, then the ghost fill).  MODE 2: the sweep of
// FFVILS::RBGS_WithFixedAverage (FFVILS.h:596-659): x^n unshifted, no ghost push, and the block
// sum of x^n (sf3d_calc_sum) into part[2][slot]; shift_push_kernel then removes the mean.
template <class T, int MODE>
__global__ __launch_bounds__(RB_NT) void rbgs_iter_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                          const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                          Coef7<T> A,
                                                          const T* __restrict__ bv, const T* __restrict__ xin,
                                                          T* __restrict__ xout, const T* __restrict__ hin,
                                                          T* __restrict__ hout, double* __restrict__ part,
                                                          size_t nslots, int n, int bcg) {
  __shared__ T s[LP * LP * LP];
  __shared__ double red[RB_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x;
  const int tx = t & 15, ty = (t >> 4) & 15, tz = t >> 8;
  const size_t base = (size_t)slot * NC;
  const size_t hbase = (size_t)slot * NH;
#pragma unroll
  for (int kk = 0; kk < 8; kk++) {
    const int k = tz * 8 + kk;
    s[lds(tx, ty, k)] = xin[base + tx + NB * (ty + NB * k)];
  }
  for (int q = t; q < NH; q += RB_NT) {
    int f = q >> 8, p = q & 255, i, j, k;
    face_cell(f, p & 15, p >> 4, -1, i, j, k);
    s[lds(i, j, k)] = hin[hbase + q];
  }
  __syncthreads();
  const bool sweep = n <= st[c].countMax;
  const T omega = (T)st[c].omega;
  const int par = (tx + ty) & 1;
  double rr = 0.0, bb = 0.0;
  T xr[4], sA[4][7], sb[4], xj[8];
#pragma unroll
  for (int m = 0; m < 4; m++) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = tz * 8 + 2 * m + h;
      const bool isred = (h == par);
      const size_t g = base + tx + NB * (ty + NB * k);
      const int li = lds(tx, ty, k);
      T a[7];
#pragma unroll
      for (int q = 0; q < 7; q++) a[q] = A.p[q][g];
      const T b0 = bv[g];
      const T xp = s[li];
      const T rs = b0 - a[1] * s[li - 1] - a[2] * s[li + 1] - a[3] * s[li - LP] - a[4] * s[li + LP] -
                   a[5] * s[li - LP * LP] - a[6] * s[li + LP * LP];
      const T r = rs - a[0] * xp;
      rr += (double)r * (double)r;
      if (n == 1) bb += (double)b0 * (double)b0;
      if (MODE == 1) {
        xj[2 * m + h] = xp - omega * (xp - rs / a[0]);
      } else if (isred) {
        xr[m] = xp - omega * (xp - rs / a[0]);
      } else {
#pragma unroll
        for (int q = 0; q < 7; q++) sA[m][q] = a[q];
        sb[m] = b0;
      }
    }
  }
  double xsum = 0.0;
  if (sweep && MODE == 1) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) s[lds(tx, ty, tz * 8 + q)] = xj[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q++) xout[base + tx + NB * (ty + NB * (tz * 8 + q))] = xj[q];
    push_halo<T, RB_NT>(B, cs[c], bcg, slot, [&](int i, int j, int k) { return s[lds(i, j, k)]; }, hin, hout);
  } else if (sweep) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; m++) s[lds(tx, ty, tz * 8 + 2 * m + par)] = xr[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int k = tz * 8 + 2 * m + (1 - par);
      const int li = lds(tx, ty, k);
      const T xp = s[li];
      const T rs = sb[m] - sA[m][1] * s[li - 1] - sA[m][2] * s[li + 1] - sA[m][3] * s[li - LP] -
                   sA[m][4] * s[li + LP] - sA[m][5] * s[li - LP * LP] - sA[m][6] * s[li + LP * LP];
      s[li] = xp - omega * (xp - rs / sA[m][0]);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      const int k = tz * 8 + kk;
      const T v = s[lds(tx, ty, k)];
      xout[base + tx + NB * (ty + NB * k)] = v;
      if (MODE == 2) xsum += (double)v;
    }
    if (MODE == 0)
      push_halo<T, RB_NT>(B, cs[c], bcg, slot, [&](int i, int j, int k) { return s[lds(i, j, k)]; }, hin, hout);
  }
  // residual of x^{n-1} (and b.b on the first launch)
  double rsum = block_sum<RB_NT>(rr, red);
  double bsum = (n == 1) ? block_sum<RB_NT>(bb, red) : 0.0;
  double ssum = MODE == 2 ? block_sum<RB_NT>(xsum, red) : 0.0;
  if (t == 0) {
    part[slot] = rsum;
    part[nslots + slot] = bsum;
    if (MODE == 2) part[2 * nslots + slot] = ssum;
  }
}

// RBGS_WithFixedAverage, second half of an iteration: x^n += a (a = -sum/Ncells from the
// finalisation, ShiftByA / adda, blas.f90:162-193), then the ghost fill of the shifted x^n.
template <class T>
__global__ __launch_bounds__(256) void shift_push_kernel(const int* __restrict__ active,
                                                         const DevBlock* __restrict__ blk,
                                                         const DevCase* __restrict__ cs,
                                                         const SolveState* __restrict__ st, T* __restrict__ x,
                                                         const T* __restrict__ hin, T* __restrict__ hout, int bcg) {
  __shared__ T s[LP * LP * LP];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const T a = (T)st[c].alpha;
  for (int k = 0; k < NB; k++) {
    const size_t g = base + tx + NB * (ty + NB * k);
    const T v = x[g] + a;
    x[g] = v;
    s[lds(tx, ty, k)] = v;
  }
  __syncthreads();
  push_halo<T, 256>(B, cs[c], bcg, slot, [&](int i, int j, int k) { return s[lds(i, j, k)]; }, hin, hout);
}

// Copy buffer 1 into buffer 0 for every case whose result landed in buffer 1.
template <class T>
__global__ __launch_bounds__(256) void select_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                     const SolveState* __restrict__ st, T* __restrict__ x0,
                                                     const T* __restrict__ x1, T* __restrict__ h0,
                                                     const T* __restrict__ h1) {
  const int slot = active[blockIdx.x];
  if (st[blk[slot].case_id].parity == 0) return;
  const size_t base = (size_t)slot * NC, hb = (size_t)slot * NH;
  for (int q = threadIdx.x; q < NC; q += 256) x0[base + q] = x1[base + q];
  for (int q = threadIdx.x; q < NH; q += 256) h0[hb + q] = h1[hb + q];
}

__global__ void solve_init_kernel(SolveState* st, int ncase, double omega, double eps, int countMax,
                                  const SolveState* outer) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncase) return;
  SolveState s = {};
  if (outer && outer[c].done) s.done = 1;  // inner solve of a case whose defect correction finished
  s.omega = omega;
  s.eps = eps;
  // a restarted inner solve gets what is left of the case's own budget
  s.countMax = outer ? countMax - outer[c].used : countMax;
  if (s.countMax <= 0) s.done = 1;
  s.count = 0;
  s.rr0 = 1.0;
  st[c] = s;
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
template <class T>
void launch_ghost_pull(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                       const T* x, T* h, int bcg, int updater, bool multilevel) {
  if (nact == 0) return;
  hipLaunchKernelGGL(ghost_pull_kernel<T>, dim3(nact * 6), dim3(256), 0, sm, active, blk, cs, x, h, bcg, updater, 0);
  if (multilevel)
    hipLaunchKernelGGL(ghost_pull_kernel<T>, dim3(nact * 6), dim3(256), 0, sm, active, blk, cs, x, h, bcg, updater, 1);
}

template <class T>
void launch_rbgs_iter(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                      SolveState* st, Coef7<T> A, const T* b, const T* xin, T* xout, const T* hin,
                      T* hout, double* part, size_t nslots, int n, int bcg, int mode) {
  if (nact == 0) return;
  if (mode == 0)
    hipLaunchKernelGGL((rbgs_iter_kernel<T, 0>), dim3(nact), dim3(RB_NT), 0, sm, active, blk, cs, st, A, b, xin,
                       xout, hin, hout, part, nslots, n, bcg);
  else if (mode == 1)
    hipLaunchKernelGGL((rbgs_iter_kernel<T, 1>), dim3(nact), dim3(RB_NT), 0, sm, active, blk, cs, st, A, b, xin,
                       xout, hin, hout, part, nslots, n, bcg);
  else
    hipLaunchKernelGGL((rbgs_iter_kernel<T, 2>), dim3(nact), dim3(RB_NT), 0, sm, active, blk, cs, st, A, b, xin,
                       xout, hin, hout, part, nslots, n, bcg);
}

template <class T>
void launch_shift_push(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                       const SolveState* st, T* x, const T* hin, T* hout, int bcg) {
  if (nact == 0) return;
  hipLaunchKernelGGL(shift_push_kernel<T>, dim3(nact), dim3(256), 0, sm, active, blk, cs, st, x, hin, hout, bcg);
}

template <class T>
void launch_select(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const SolveState* st, T* x0,
                   const T* x1, T* h0, const T* h1) {
  if (nact == 0) return;
  hipLaunchKernelGGL(select_kernel<T>, dim3(nact), dim3(256), 0, sm, active, blk, st, x0, x1, h0, h1);
}

__global__ void add_used_kernel(SolveState* outer, const SolveState* inner, int ncase) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncase) outer[c].used += inner[c].count;
}

void launch_add_used(hipStream_t sm, SolveState* outer, const SolveState* inner, int ncase) {
  if (ncase > 0) hipLaunchKernelGGL(add_used_kernel, dim3((ncase + 63) / 64), dim3(64), 0, sm, outer, inner, ncase);
}

void launch_solve_init(hipStream_t sm, SolveState* st, int ncase, double omega, double eps, int countMax,
                       const SolveState* outer) {
  if (ncase <= 0) return;
  hipLaunchKernelGGL(solve_init_kernel, dim3((ncase + 63) / 64), dim3(64), 0, sm, st, ncase, omega, eps, countMax,
                     outer);
}

template void launch_ghost_pull<double>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, const double*,
                                        double*, int, int, bool);
template void launch_ghost_pull<float>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, const float*,
                                       float*, int, int, bool);
template void launch_rbgs_iter<double>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, SolveState*,
                                       Coef7<double>, const double*, const double*, double*, const double*, double*,
                                       double*, size_t, int, int, int);
template void launch_rbgs_iter<float>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, SolveState*,
                                      Coef7<float>, const float*, const float*, float*, const float*, float*, double*,
                                      size_t, int, int, int);
template void launch_shift_push<double>(hipStream_t, const int*, int, const DevBlock*, const DevCase*,
                                        const SolveState*, double*, const double*, double*, int);
template void launch_shift_push<float>(hipStream_t, const int*, int, const DevBlock*, const DevCase*,
                                       const SolveState*, float*, const float*, float*, int);
template void launch_select<double>(hipStream_t, const int*, int, const DevBlock*, const SolveState*, double*,
                                    const double*, double*, const double*);
template void launch_select<float>(hipStream_t, const int*, int, const DevBlock*, const SolveState*, float*,
                                   const float*, float*, const float*);

}  // namespace ffvhc

namespace ffvhc {

// FFVILS::Fill / Copy over the selected blocks (interior and ghost layer)
template <class T>
__global__ __launch_bounds__(256) void fill_kernel(const int* __restrict__ active, T* __restrict__ x, T* __restrict__ h,
                                                   T v) {
  const int slot = active[blockIdx.x];
  for (int q = threadIdx.x; q < NC; q += 256) x[(size_t)slot * NC + q] = v;
  for (int q = threadIdx.x; q < NH; q += 256) h[(size_t)slot * NH + q] = v;
}

template <class T>
__global__ __launch_bounds__(256) void copy_kernel(const int* __restrict__ active, T* __restrict__ y, T* __restrict__ hy,
                                                   const T* __restrict__ x, const T* __restrict__ hx) {
  const int slot = active[blockIdx.x];
  for (int q = threadIdx.x; q < NC; q += 256) y[(size_t)slot * NC + q] = x[(size_t)slot * NC + q];
  for (int q = threadIdx.x; q < NH; q += 256) hy[(size_t)slot * NH + q] = hx[(size_t)slot * NH + q];
}

template <class T>
void launch_fill(hipStream_t sm, const int* active, int nact, T* x, T* h, T v) {
  if (nact) hipLaunchKernelGGL(fill_kernel<T>, dim3(nact), dim3(256), 0, sm, active, x, h, v);
}
template <class T>
void launch_copy(hipStream_t sm, const int* active, int nact, T* y, T* hy, const T* x, const T* hx) {
  if (nact) hipLaunchKernelGGL(copy_kernel<T>, dim3(nact), dim3(256), 0, sm, active, y, hy, x, hx);
}
template void launch_fill<double>(hipStream_t, const int*, int, double*, double*, double);
template void launch_fill<float>(hipStream_t, const int*, int, float*, float*, float);
template void launch_copy<double>(hipStream_t, const int*, int, double*, double*, const double*, const double*);
template void launch_copy<float>(hipStream_t, const int*, int, float*, float*, const float*, const float*);

}  // namespace ffvhc
