// Coarse (one value per block) correction of the block preconditioners: the coarse operator
// A_c = P^T A P with P the block indicator, its dense inverse per case, and its application.
//
// The preconditioned direction of the two-level methods is  z = z_loc(v) + P A_c^-1 P^T v,
// z_loc the block-local SOR sweeps.  A ghost value of P e is the source block's e for every
// face kind (same-level copy, the fine-to-coarse mean and the coarse-to-fine trilinear
// interpolation, whose weights sum to one, and the homogeneous Neumann outer ghost of the
// preconditioned vectors), so the matvec adds e of the source block to each staged ghost and the
// corrected vector is never stored.
#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

constexpr int CR_NT = 256;

// Row of A_c for block `slot`: rows[slot][0] = coupling to itself, rows[slot][1 + 4 f + s] =
// coupling to nbr[f][s] (sum of the band-q coefficients of the face cells whose ghost comes from
// that block).  Sums in a fixed order: deterministic.
__global__ __launch_bounds__(CR_NT) void coarse_rows_kernel(const int* __restrict__ active,
                                                            const DevBlock* __restrict__ blk, Coef7<double> A,
                                                            double* __restrict__ rows) {
  __shared__ double red[CR_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int t = threadIdx.x;
  const size_t base = (size_t)slot * NC;
  // all coefficients of the block, then the face couplings that leave it
  double all = 0.0;
  for (int q = t; q < NC; q += CR_NT) {
    double s = A.p[0][base + q];
#pragma unroll
    for (int b = 1; b < 7; b++) s += A.p[b][base + q];
    all += s;
  }
  const double tot = block_sum<CR_NT>(all, red);
  double self = tot;  // valid in thread 0
  for (int f = 0; f < 6; f++) {
    const int a = t & 15, b = t >> 4;  // one face position per thread
    int i, j, k;
    face_cell(f, a, b, 0, i, j, k);
    const double c = A.p[1 + f][base + i + NB * (j + NB * k)];
    const int sub = B.ftype[f] == FACE_FINER ? [&] {
      int u0, u1;
      face_ab_to_o(f, a, b, u0, u1);
      return (u0 >= NB / 2 ? 1 : 0) + (u1 >= NB / 2 ? 2 : 0);
    }()
                                             : 0;
    for (int s = 0; s < 4; s++) {
      const double v = block_sum<CR_NT>(sub == s ? c : 0.0, red);
      if (t == 0) {
        rows[(size_t)slot * 25 + 1 + 4 * f + s] = B.ftype[f] == FACE_OUTER ? 0.0 : v;
        if (B.ftype[f] != FACE_OUTER) self -= v;
      }
    }
  }
  if (t == 0) rows[(size_t)slot * 25] = self;
}

void launch_coarse_rows(hipStream_t sm, const int* active, int nact, const DevBlock* blk, Coef7<double> A,
                        double* rows) {
  if (nact) hipLaunchKernelGGL(coarse_rows_kernel, dim3(nact), dim3(CR_NT), 0, sm, active, blk, A, rows);
}

// In-place Gauss-Jordan inversion of a dense n x n matrix (row-major), one launch per pivot k.
// snap holds the pivot row and column of step k in buffer (k & 1): each launch reads only the
// snapshot of its step and its own element, and the threads owning row / column k + 1 write the
// next snapshot, so no element is read after another thread of the launch wrote it.
//   snap layout per buffer: row[n], col[n]
__global__ void gj_snap_kernel(const double* __restrict__ M, int n, double* __restrict__ snap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  snap[i] = M[i];                // row 0
  snap[n + i] = M[(size_t)i * n];  // column 0
}

__global__ void gj_step_kernel(double* __restrict__ M, int n, int k, double* __restrict__ snap) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  const double* cur = snap + (size_t)(k & 1) * 2 * n;
  double* nxt = snap + (size_t)((k + 1) & 1) * 2 * n;
  const double p = cur[k];  // row[k] = M[k][k]
  double v;
  if (i == k && j == k) v = 1.0 / p;
  else if (i == k) v = cur[j] / p;
  else if (j == k) v = -cur[n + i] / p;
  else v = M[e] - cur[n + i] * cur[j] / p;
  M[e] = v;
  if (i == k + 1) nxt[j] = v;
  if (j == k + 1) nxt[n + i] = v;
}

void gauss_jordan_inverse(hipStream_t sm, double* M, int n, double* snap) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gj_snap_kernel, dim3((n + 255) / 256), dim3(256), 0, sm, (const double*)M, n, snap);
  const size_t nn = (size_t)n * n;
  const dim3 g((unsigned)((nn + 255) / 256));
  for (int k = 0; k < n; k++) hipLaunchKernelGGL(gj_step_kernel, g, dim3(256), 0, sm, M, n, k, snap);
}

__global__ void to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = (float)x[e];
}

void launch_to_f32(hipStream_t sm, const double* x, float* y, size_t n) {
  if (n) hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, sm, x, y, n);
}

// e[slot] = sum_b inv[a][b] c[b] for every block a of every selected case: one wavefront per row,
// lanes strided over b, then a fixed shuffle tree (deterministic).  c[b] = part[slot_begin + b],
// the block sums of the vector being preconditioned; inv is the case's FP32 inverse (row-major,
// at offset off[case]).
__global__ __launch_bounds__(256) void coarse_apply_kernel(const int* __restrict__ active, int nact,
                                                           const DevBlock* __restrict__ blk,
                                                           const DevCase* __restrict__ cs,
                                                           const SolveState* __restrict__ st,
                                                           const float* __restrict__ inv,
                                                           const long long* __restrict__ off,
                                                           const double* __restrict__ c, double* __restrict__ e) {
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (w >= nact) return;
  const int slot = active[w];
  const int cid = blk[slot].case_id;
  if (st[cid].done) return;
  const DevCase& C = cs[cid];
  const int n = C.nblk, a = slot - C.slot_begin;
  const float* row = inv + off[cid] + (size_t)a * n;
  double s = 0.0;
  for (int b = lane; b < n; b += 64) s += (double)row[b] * c[C.slot_begin + b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) e[slot] = s;
}

void launch_coarse_apply(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                         const SolveState* st, const float* inv, const long long* off, const double* c, double* e) {
  if (nact)
    hipLaunchKernelGGL(coarse_apply_kernel, dim3((nact + 3) / 4), dim3(256), 0, sm, active, nact, blk, cs, st, inv,
                       off, c, e);
}

}  // namespace ffvhc
