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

// Dense A_c of every selected case from the rows: M (zeroed before) gets, in the row of block a,
// the self coupling on the diagonal and each face coupling in its neighbour's column.  One thread
// per block writes its whole row (repeated neighbours add up in order).
__global__ void coarse_scatter_kernel(const int* __restrict__ active, int nact, const DevBlock* __restrict__ blk,
                                      const DevCase* __restrict__ cs, const long long* __restrict__ off,
                                      const double* __restrict__ rows, double* __restrict__ M) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nact) return;
  const int slot = active[w];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int n = C.nblk, b0 = C.slot_begin;
  double* row = M + off[B.case_id] + (size_t)(slot - b0) * n;
  const double* r = rows + (size_t)slot * 25;
  row[slot - b0] += r[0];
  for (int f = 0; f < 6; f++) {
    if (B.ftype[f] == FACE_OUTER) continue;
    const int ns = B.ftype[f] == FACE_FINER ? 4 : 1;
    for (int s = 0; s < ns; s++) row[B.nbr[f][s] - b0] += r[1 + 4 * f + s];
  }
}

// Decouple the last coarse unknown of a singular (all-Neumann) case: zero its row and column
// outside the diagonal.
__global__ void coarse_decouple_kernel(double* __restrict__ M, int n) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n - 1) return;
  M[(size_t)(n - 1) * n + b] = 0.0;
  M[(size_t)b * n + (n - 1)] = 0.0;
}

__global__ void zero_kernel(double* __restrict__ x, size_t n) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[e] = 0.0;
}

void launch_coarse_assemble(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                            const long long* off, const double* rows, double* M, size_t total) {
  if (!nact) return;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, sm, M, total);
  hipLaunchKernelGGL(coarse_scatter_kernel, dim3((nact + 255) / 256), dim3(256), 0, sm, active, nact, blk, cs, off,
                     rows, M);
}

void launch_coarse_decouple(hipStream_t sm, double* M, int n) {
  if (n > 1) hipLaunchKernelGGL(coarse_decouple_kernel, dim3((n + 254) / 256), dim3(256), 0, sm, M, n);
}

// In-place blocked Gauss-Jordan inversion of a dense n x n matrix (row-major, no pivoting: the
// coarse operators are diagonally dominant).  For each panel P of GJB pivots, with D = M_PP:
//   M_PP <- D^-1,  M_Pj <- D^-1 M_Pj,  M_iP <- -M_iP D^-1,  M_ij <- M_ij - M_iP D^-1 M_Pj
// (i, j outside P) -- the scalar Gauss-Jordan step applied to a block of pivots.  Three launches
// per panel (the trailing update also stores the new panel rows and inverts the next panel's
// diagonal block; the column step works in place within a workgroup); no launch reads a value
// another workgroup of it writes.
constexpr int GJB = 32;

// Unblocked Gauss-Jordan inversion of the leading kb x kb part of d in place; one thread per entry
// (r, c) of a GJB * GJB workgroup
__device__ inline void gj_inplace(double (*d)[GJB + 1], int r, int c, int kb) {
  const bool in = r < kb && c < kb;
  for (int k = 0; k < kb; k++) {
    double v = 0.0;
    if (in) {
      const double p = d[k][k], rk = d[k][c], ck = d[r][k];
      if (r == k && c == k) v = 1.0 / p;
      else if (r == k) v = rk / p;
      else if (c == k) v = -ck / p;
      else v = d[r][c] - ck * rk / p;
    }
    __syncthreads();
    if (in) d[r][c] = v;
    __syncthreads();
  }
}

// D <- D^-1 for the first panel (later panels' diagonal blocks are inverted by the update kernel)
__global__ __launch_bounds__(GJB * GJB) void gjb_diag_kernel(double* __restrict__ M, int n, int k0, int kb) {
  __shared__ double d[GJB][GJB + 1];
  const int r = threadIdx.x / GJB, c = threadIdx.x % GJB;
  const bool in = r < kb && c < kb;
  if (in) d[r][c] = M[(size_t)(k0 + r) * n + k0 + c];
  __syncthreads();
  gj_inplace(d, r, c, kb);
  if (in) M[(size_t)(k0 + r) * n + k0 + c] = d[r][c];
}

// T[r][j] = (D^-1 M_P)[r][j] for the columns j outside the panel
__global__ __launch_bounds__(256) void gjb_row_kernel(const double* __restrict__ M, int n, int k0, int kb,
                                                      double* __restrict__ T) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (j >= n || (j >= k0 && j < k0 + kb)) return;
  const double* dinv = M + (size_t)(k0 + r) * n + k0;
  double s = 0.0;
  for (int c = 0; c < kb; c++) s += dinv[c] * M[(size_t)(k0 + c) * n + j];
  T[(size_t)r * n + j] = s;
}

// M_ij -= sum_c M_i,k0+c T[c][j] for i, j outside the panel, and the panel rows M_Pj <- T[.][j]
// (j outside the panel: nothing here reads them); GJB x GJB tiles, one entry per thread.  The
// workgroup whose tile is the next panel's diagonal block (final once this update is done: the
// column step touches only this panel's columns) inverts it in LDS right away, so the next panel
// needs no diagonal launch of its own.
__global__ __launch_bounds__(GJB * GJB) void gjb_update_kernel(double* __restrict__ M, int n, int k0, int kb,
                                                               const double* __restrict__ T) {
  __shared__ double a[GJB][GJB + 1], t[GJB][GJB + 1];
  const int r = threadIdx.x / GJB, cc = threadIdx.x % GJB;
  // workgroup 0 takes the next panel's diagonal tile (its inversion is the longest job, so it
  // starts first) and that tile's workgroup takes tile (0, 0)
  const int tn = (k0 + GJB) / GJB;
  int bx = blockIdx.x, by = blockIdx.y;
  if (tn < (int)gridDim.x) {
    if (bx == 0 && by == 0) bx = by = tn;
    else if (bx == tn && by == tn) bx = by = 0;
  }
  const int i0 = by * GJB, j0 = bx * GJB, i = i0 + r, j = j0 + cc;
  a[r][cc] = (i < n && cc < kb) ? M[(size_t)i * n + k0 + cc] : 0.0;
  t[r][cc] = (j < n && r < kb) ? T[(size_t)r * n + j] : 0.0;
  __syncthreads();
  double v = 0.0;
  if (i < n && j < n && !(j >= k0 && j < k0 + kb)) {
    if (i >= k0 && i < k0 + kb) {
      v = T[(size_t)(i - k0) * n + j];
    } else {
      double s = 0.0;
#pragma unroll 8
      for (int c = 0; c < GJB; c++) s += a[r][c] * t[c][cc];
      v = M[(size_t)i * n + j] - s;
    }
    M[(size_t)i * n + j] = v;
  }
  const int k1 = k0 + GJB;
  if (i0 != k1 || j0 != k1) return;  // uniform over the workgroup
  const int kb1 = n - k1 < GJB ? n - k1 : GJB;
  const bool in = r < kb1 && cc < kb1;
  __syncthreads();  // every read of a and t is done
  a[r][cc] = in ? v : 0.0;
  __syncthreads();
  gj_inplace(a, r, cc, kb1);
  if (in) M[(size_t)(k1 + r) * n + k1 + cc] = a[r][cc];
}
// M_iP <- -M_iP D^-1 for the rows outside the panel, in place: the GJB threads of a row share a
// workgroup (256 / GJB rows each), read the row's panel values, synchronise, then write them
__global__ __launch_bounds__(256) void gjb_col_kernel(double* __restrict__ M, int n, int k0, int kb) {
  const int c2 = threadIdx.x % GJB, i = blockIdx.x * (256 / GJB) + threadIdx.x / GJB;
  const bool act = i < n && c2 < kb && !(i >= k0 && i < k0 + kb);
  double s = 0.0;
  if (act) {
    const double* a = M + (size_t)i * n + k0;
    for (int c = 0; c < kb; c++) s += a[c] * M[(size_t)(k0 + c) * n + k0 + c2];
  }
  __syncthreads();
  if (act) M[(size_t)i * n + k0 + c2] = -s;
}

// Helpers of the wide-panel Gauss-Jordan (gauss_jordan_inverse_gemm in solver.cpp), whose three
// products per panel run as FP64 GEMMs.  All matrices row-major; the panel K = [k0, k0 + kb).
// D (kb x kb, pitch kb) <- M_KK
__global__ __launch_bounds__(256) void gjw_get_diag_kernel(const double* __restrict__ M, int n, int k0, int kb,
                                                           double* __restrict__ D) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= kb * kb) return;
  const int r = e / kb, c = e % kb;
  D[e] = M[(size_t)(k0 + r) * n + k0 + c];
}
// C (n x kb, pitch kb) <- M_iK for the rows i outside K, 0 for the rows in K
__global__ __launch_bounds__(256) void gjw_get_cols_kernel(const double* __restrict__ M, int n, int k0, int kb,
                                                           double* __restrict__ C) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * kb) return;
  const int i = (int)(e / kb), c = (int)(e % kb);
  C[e] = (i >= k0 && i < k0 + kb) ? 0.0 : M[(size_t)i * n + k0 + c];
}
// T (kb x n, pitch n): the columns in K set to 0 (T = D^-1 M_K. before this)
__global__ __launch_bounds__(256) void gjw_zero_tcols_kernel(double* __restrict__ T, int n, int k0, int kb) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= kb * kb) return;
  T[(size_t)(e / kb) * n + k0 + e % kb] = 0.0;
}
// M_K. <- T outside K and D^-1 (in D) inside K
__global__ __launch_bounds__(256) void gjw_put_rows_kernel(double* __restrict__ M, int n, int k0, int kb,
                                                           const double* __restrict__ T,
                                                           const double* __restrict__ D) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * kb) return;
  const int r = (int)(e / n), j = (int)(e % n);
  M[(size_t)(k0 + r) * n + j] = (j >= k0 && j < k0 + kb) ? D[(size_t)r * kb + j - k0] : T[e];
}

void launch_gjw_get_diag(hipStream_t sm, const double* M, int n, int k0, int kb, double* D) {
  hipLaunchKernelGGL(gjw_get_diag_kernel, dim3((unsigned)((kb * kb + 255) / 256)), dim3(256), 0, sm, M, n, k0, kb, D);
}
void launch_gjw_get_cols(hipStream_t sm, const double* M, int n, int k0, int kb, double* C) {
  hipLaunchKernelGGL(gjw_get_cols_kernel, dim3((unsigned)(((size_t)n * kb + 255) / 256)), dim3(256), 0, sm, M, n, k0,
                     kb, C);
}
void launch_gjw_zero_tcols(hipStream_t sm, double* T, int n, int k0, int kb) {
  hipLaunchKernelGGL(gjw_zero_tcols_kernel, dim3((unsigned)((kb * kb + 255) / 256)), dim3(256), 0, sm, T, n, k0, kb);
}
void launch_gjw_put_rows(hipStream_t sm, double* M, int n, int k0, int kb, const double* T, const double* D) {
  hipLaunchKernelGGL(gjw_put_rows_kernel, dim3((unsigned)(((size_t)n * kb + 255) / 256)), dim3(256), 0, sm, M, n, k0,
                     kb, T, D);
}

// T: scratch of GJB * n doubles.  Order per panel: D^-1; T = D^-1 M_P; the trailing update
// (old M_iP, T) with panel rows <- T; M_iP <- -M_iP D^-1.
void gauss_jordan_inverse(hipStream_t sm, double* M, int n, double* T) {
  if (n <= 0) return;
  const unsigned nt = (unsigned)((n + 255) / 256), ntt = (unsigned)((n + GJB - 1) / GJB);
  const unsigned nr = (unsigned)((n + 256 / GJB - 1) / (256 / GJB));
  hipLaunchKernelGGL(gjb_diag_kernel, dim3(1), dim3(GJB * GJB), 0, sm, M, n, 0, n < GJB ? n : GJB);
  for (int k0 = 0; k0 < n; k0 += GJB) {
    const int kb = n - k0 < GJB ? n - k0 : GJB;
    hipLaunchKernelGGL(gjb_row_kernel, dim3(nt, kb), dim3(256), 0, sm, (const double*)M, n, k0, kb, T);
    hipLaunchKernelGGL(gjb_update_kernel, dim3(ntt, ntt), dim3(GJB * GJB), 0, sm, M, n, k0, kb, (const double*)T);
    hipLaunchKernelGGL(gjb_col_kernel, dim3(nr), dim3(256), 0, sm, M, n, k0, kb);
  }
}

// FP32 copy of a case's FP64 inverse with the row pitch padded to np (a multiple of 4, pad zero),
// so that every row starts 16-byte aligned for float4 loads
__global__ void to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, int n, int np) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * np) return;
  const int a = (int)(e / np), b = (int)(e % np);
  y[e] = b < n ? (float)x[(size_t)a * n + b] : 0.0f;
}

void launch_to_f32(hipStream_t sm, const double* x, float* y, int n, int np) {
  const size_t total = (size_t)n * np;
  if (total) hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, sm, x, y, n, np);
}

__device__ inline int coarse_pitch(int n) { return (n + 3) & ~3; }

// e = A_c^-1 c for each active block's row: one wavefront per row, the row read as float4
// (pitch coarse_pitch(n), offsets off[case]); c(b) gives the coarse right-hand side of block b
// (0 past n); the sum order is fixed (lane-strided float4 groups, then the wave reduction), so a
// case's result does not depend on what else is batched with it
template <class Rhs>
__device__ inline double coarse_row_dot(const float* __restrict__ row, int n, int lane, const Rhs& c) {
  const float4* r4 = reinterpret_cast<const float4*>(row);
  double s = 0.0;
  const int nq = coarse_pitch(n) / 4;
#pragma unroll 4
  for (int q = lane; q < nq; q += 64) {
    const float4 v = r4[q];
    const int b = 4 * q;
    s += (double)v.x * c(b);
    s += (double)v.y * c(b + 1);
    s += (double)v.z * c(b + 2);
    s += (double)v.w * c(b + 3);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// The same sum with the row's first CA_PRE float4 per lane already loaded (pre[i] = r4[lane + 64 i])
constexpr int CA_PRE = 6;
template <class Rhs>
__device__ inline double coarse_row_dot_pre(const float* __restrict__ row, int n, int lane, const float4* pre,
                                            const Rhs& c) {
  const float4* r4 = reinterpret_cast<const float4*>(row);
  double s = 0.0;
  const int nq = coarse_pitch(n) / 4;
#pragma unroll
  for (int i = 0; i < CA_PRE; i++) {
    const int q = lane + 64 * i;
    if (q < nq) {
      const float4 v = pre[i];
      const int b = 4 * q;
      s += (double)v.x * c(b);
      s += (double)v.y * c(b + 1);
      s += (double)v.z * c(b + 2);
      s += (double)v.w * c(b + 3);
    }
  }
#pragma unroll 4
  for (int q = lane + 64 * CA_PRE; q < nq; q += 64) {
    const float4 v = r4[q];
    const int b = 4 * q;
    s += (double)v.x * c(b);
    s += (double)v.y * c(b + 1);
    s += (double)v.z * c(b + 2);
    s += (double)v.w * c(b + 3);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}
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
  const float* row = inv + off[cid] + (size_t)a * coarse_pitch(n);
  const double* cc = c + C.slot_begin;
  const double s = coarse_row_dot(row, n, lane, [&](int b) { return b < n ? cc[b] : 0.0; });
  if (lane == 0) e[slot] = s;
}

// Multiplicative two-level methods: e = A_c^-1 P^T v with P^T v formed from the block sums of r,
// q_ and p the previous kernels kept (ce[2], ce[3], ce[4] per slot) and the case scalars:
// p = r (first iteration) or r + beta (p - gamma q_), s = r - alpha q_.  A workgroup of CA_NT
// threads handles CA_NT / 64 rows of one case; it forms c once in LDS (up to CA_MAXN blocks per
// case) when all its rows belong to one case, else each wave forms c on the fly.
constexpr int CA_NT = 256, CA_MAXN = 4096;

struct CoarseRhs {
  const double *br, *bq, *bp;
  bool first;
  double beta, ng, na;
  int mode;
  // mode 1 takes alpha from the caller (formed in the kernel); mode 0 reads its scalars from S,
  // or (fused, the previous iteration's finalisation formed in the kernel) takes beta from the
  // caller and tells the first iteration by count, S.iter and S.beta being rewritten meanwhile
  __device__ CoarseRhs(const double* ce, size_t nslots, const DevCase& C, const SolveState& S, int m, double alpha,
                       bool fused = false, double fbeta = 0.0, int count = 0)
      : br(ce + 2 * nslots + C.slot_begin), bq(ce + 3 * nslots + C.slot_begin), bp(ce + 4 * nslots + C.slot_begin),
        first(m == 0 && ((fused ? count == 1 : S.iter == 0) || S.restart)),
        beta(m == 0 ? (fused ? fbeta : S.beta) : 0.0), ng(m == 0 ? -S.gamma : 0.0), na(-alpha), mode(m) {}
  __device__ double operator()(int b) const {
    return mode == 0 ? (first ? br[b] : br[b] + beta * (bp[b] + ng * bq[b])) : na * bq[b] + br[b];
  }
};

// mode 1: apart holds the Q stage's block partials (q_.r0); every wave forms its case's
// alpha = rho / (q_.r0) (FFVILS.h:868-872) itself with a fixed-order wave sum (the same value in
// every wave of the case, whatever else the launch holds) and the wave of the case's first block
// stores it, so no alpha finalisation launch is needed.
// mode 0 with xpart (the XR partials of iteration count - 1, r.r and r.r0): every wave forms that
// iteration's finalisation (IsConverged, the next rho and beta) itself, with the per-case
// finalisation's sums in its order (the same bits), from fields of the case state that nothing
// rewrites during this launch (rho_n in rr0, put there by XR); the wave of the case's first block
// stores it as the finalisation would, so no finalisation launch sits between XR and this one.  A
// case that converged there takes no correction.
__global__ __launch_bounds__(CA_NT) void coarse_apply_mul_kernel(const int* __restrict__ active, int nact,
                                                                 const DevBlock* __restrict__ blk,
                                                                 const DevCase* __restrict__ cs,
                                                                 SolveState* __restrict__ st,
                                                                 const float* __restrict__ inv,
                                                                 const long long* __restrict__ off,
                                                                 double* __restrict__ ce, size_t nslots, int mode,
                                                                 const double* __restrict__ apart, int count,
                                                                 const double* __restrict__ xpart) {
  __shared__ double cl[CA_MAXN];
  __shared__ double fred[2 * (CA_NT / 64)], fsum[2];
  constexpr int R = CA_NT / 64;  // rows per workgroup
  const int first_row = blockIdx.x * R;
  const int last_row = min(first_row + R - 1, nact - 1);
  // the active list holds each case's blocks contiguously: a workgroup whose first and last rows
  // share a case stages that case's c once in LDS
  const int c0 = blk[active[first_row]].case_id;
  const int n0 = cs[c0].nblk;
  const bool same = c0 == blk[active[last_row]].case_id && coarse_pitch(n0) <= CA_MAXN;  // uniform per workgroup
  const bool uni = same && !st[c0].done;
  const int w = first_row + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  // this wave's row: its first float4s are loaded before c is staged, so the two memory round
  // trips overlap
  const int slot = active[w < nact ? w : first_row];  // a wave past the list reads a valid entry
  const int cid = blk[slot].case_id;
  const DevCase& C = cs[cid];
  const int n = C.nblk, a = slot - C.slot_begin;
  const float* row = inv + off[cid] + (size_t)a * coarse_pitch(n);
  float4 pre[CA_PRE];
  if (uni && w < nact) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
    const int nq = coarse_pitch(n) / 4;
#pragma unroll
    for (int i = 0; i < CA_PRE; i++) {
      const int q = lane + 64 * i;
      pre[i] = q < nq ? r4[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  // mode 1: alpha of the case this wave stages for (c0 in a one-case workgroup) or computes
  const int acid = uni ? c0 : cid;
  double alpha = 0.0;
  if (mode == 1 && (uni || w < nact) && !st[acid].done) {
    const DevCase& A = cs[acid];
    double v = 0.0;
#pragma unroll 8
    for (int q = lane; q < A.nblk; q += 64) v += apart[A.slot_begin + q];  // loads issued 8 at a time
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    alpha = st[acid].rr1 / v;
  }
  // mode 0 with xpart: the previous iteration's finalisation.  A one-case workgroup forms the sums
  // with all its threads as the finalisation does (its barriers taken by every thread, whatever the
  // case's state); a workgroup that straddles cases forms them per wave.
  const bool fused = mode == 0 && xpart;
  double fbeta = 0.0;
  bool fdone = false;
  if (fused) {
    double t0 = 0.0, t1 = 0.0;
    if (same) {
      case_sum2<CA_NT>(xpart, xpart + nslots, cs[c0].slot_begin, n0, fred, t0, t1);
      if (threadIdx.x == 0) {
        fsum[0] = t0;
        fsum[1] = t1;
      }
      __syncthreads();
      t0 = fsum[0];
      t1 = fsum[1];
    } else if (w < nact) {
      case_sum2_wave(xpart, xpart + nslots, cs[acid].slot_begin, cs[acid].nblk, t0, t1);
    }
    if ((same || w < nact) && !st[acid].done) {
      const double rho_n = st[acid].rr0;
      const FinXr f = fin_bi_xr_eval(st[acid], t0, t1, count - 1, rho_n);
      fbeta = f.beta;
      fdone = f.done != 0;
      if (w < nact && slot == cs[acid].slot_begin && lane == 0) fin_bi_xr_store(st[acid], f, count - 1, rho_n);
    }
  }
  if (uni && !fdone) {
    const CoarseRhs cb(ce, nslots, cs[c0], st[c0], mode, alpha, fused, fbeta, count);
    for (int b = threadIdx.x; b < coarse_pitch(n0); b += CA_NT) cl[b] = b < n0 ? cb(b) : 0.0;
  }
  __syncthreads();
  if (w >= nact || fdone) return;
  SolveState& S = st[cid];
  if (S.done) return;
  if (mode == 1 && a == 0 && lane == 0) {  // the case's first block stores alpha (FIN_BI_ALPHA)
    S.alpha = alpha;
    S.restart = 0;
  }
  double s;
  if (uni) {
    s = coarse_row_dot_pre(row, n, lane, pre, [&](int b) { return cl[b]; });
  } else {
    const CoarseRhs cb(ce, nslots, C, S, mode, alpha, fused, fbeta, count);
    s = coarse_row_dot(row, n, lane, [&](int b) { return b < n ? cb(b) : 0.0; });
  }
  if (lane == 0) ce[mode * nslots + slot] = s;
}

void launch_coarse_apply_mul(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                             SolveState* st, const float* inv, const long long* off, double* ce, size_t nslots,
                             int mode, const double* apart, int count, const double* xpart) {
  if (nact)
    hipLaunchKernelGGL(coarse_apply_mul_kernel, dim3((nact + CA_NT / 64 - 1) / (CA_NT / 64)), dim3(CA_NT), 0, sm,
                       active, nact, blk, cs, st, inv, off, ce, nslots, mode, apart, count, xpart);
}

void launch_coarse_apply(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                         const SolveState* st, const float* inv, const long long* off, const double* c, double* e) {
  if (nact)
    hipLaunchKernelGGL(coarse_apply_kernel, dim3((nact + 3) / 4), dim3(256), 0, sm, active, nact, blk, cs, st, inv,
                       off, c, e);
}

}  // namespace ffvhc
