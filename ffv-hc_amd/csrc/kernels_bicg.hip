// Jacobi right-preconditioned BiCGSTAB (FFVILS::BiCGSTAB, FFVILS.h:793-974) as five fused kernels
// per iteration; every scalar (alpha, gamma, beta, the convergence test) is computed on the device
// by the per-case finalisation (kernels_reduce.hip).  The preconditioner is one damped-Jacobi sweep from zero
// (countPreConditioner = 1), value-identical to omega * (v / Ap) (see kernels_cg.hip).
//
//   init : r = b - A x, r0 = r; b.b, r.r0                               [FFVILS.h:810-826]
//   P    : p = r | r + beta (p - gamma q_); p_ = M p, push p_ ghosts   [828-858]
//   Q    : q_ = A p_; q_.r0 -> alpha                                   [859-872]
//   S    : s = r - alpha q_; s_ = M s, push s_ ghosts                  [873-893]
//   T    : t_ = A s_; t_.s, t_.t_ -> gamma                             [894-915]
//   XR   : x += alpha p_ + gamma s_; r = s - gamma t_; r.r -> IsConverged; r.r0 -> beta [916-965]
#include <type_traits>

#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

constexpr int BI_NT = 256;

// Seeded shadow residual for the mixed-precision solver: r0 = uniform(-1, 1) from a hash of the
// block's index within its case and the cell (so a case draws the same r0 whatever slots it
// holds).  With r0 = r (the reference's choice) rho = r.r0 decays far below FP32 resolution on
// the coarse/fine octree operator; a random shadow keeps it near |r||r0|/sqrt(N).
__device__ inline float shadow_value(uint32_t blk, uint32_t cell) {
  uint32_t h = blk * 0x9E3779B1u ^ (cell + 0x7F4A7C15u) * 0x85EBCA77u;  // integer mix (murmur-style)
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return (float)((double)h * (2.0 / 4294967296.0) - 1.0);
}


// HR: the shadow residual r0 is the seeded hash vector (shadow_value), computed where it is read
// instead of stored
template <class T, bool HR = false>
__global__ __launch_bounds__(BI_NT) void bi_init_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                        const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                        Coef7<T> A, const T* __restrict__ bv, const T* __restrict__ x,
                                                        const T* __restrict__ hx, BiVecs<T> V,
                                                        double* __restrict__ part, size_t nslots,
                                                        double* __restrict__ bsr) {
  __shared__ T s[LP * LP * LP];
  __shared__ double red[BI_NT / 64];
  const int slot = active[blockIdx.x];
  const int c = blk[slot].case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const uint32_t local = (uint32_t)(slot - cs[c].slot_begin);
  constexpr int KB = Batch<T>::K;
  double bb = 0.0, rr1 = 0.0, rs = 0.0;
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
      stage_block<T, BI_NT>(s, x, hx, slot);
      __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      const int li = lds(tx, ty, k0 + kk);
      const T r = b0[kk] - a[kk][1] * s[li - 1] - a[kk][2] * s[li + 1] - a[kk][3] * s[li - LP] -
                  a[kk][4] * s[li + LP] - a[kk][5] * s[li - LP * LP] - a[kk][6] * s[li + LP * LP] -
                  a[kk][0] * s[li];
      V.r[g] = r;
      rs += (double)r;
      bb += (double)b0[kk] * (double)b0[kk];
      if (HR) {
        rr1 += (double)r * (double)shadow_value(local, (uint32_t)(g - base));
      } else {
        V.r0[g] = r;
        rr1 += (double)r * (double)r;
      }
    }
  }
  double sb = block_sum<BI_NT>(bb, red);
  double sr = block_sum<BI_NT>(rr1, red);
  if (t == 0) {
    part[slot] = sb;
    part[nslots + slot] = sr;
  }
  if (bsr) {  // block sums of r for the multiplicative coarse correction
    const double v = block_sum<BI_NT>(rs, red);
    if (t == 0) bsr[slot] = v;
  }
}

// P: p update and p_ = M p with its ghost layer (mode 0); S: s = r - alpha q_, s_ = M s (mode 1)
// TA: storage precision of the matrix (FP32 copies for an FP64 solve are widened on load)
template <class T, int MODE, class TA = T>
__global__ __launch_bounds__(BI_NT) void bi_precond_kernel(const int* __restrict__ active,
                                                           const DevBlock* __restrict__ blk,
                                                           const DevCase* __restrict__ cs,
                                                           const SolveState* __restrict__ st, Coef7<TA> A, BiVecs<T> V,
                                                           int count) {
  __shared__ T s[LP * LP * LP];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const T omega = (T)st[c].omega;
  const T ng = (T)(-st[c].gamma), beta = (T)st[c].beta, na = (T)(-st[c].alpha);
  // count 0: the iteration number comes from the case state (a replayed graph has no per-launch
  // arguments); FIN_BI_XR of the previous iteration stored it in iter
  const bool first = (count > 0 ? count : st[c].iter + 1) == 1 || st[c].restart;
  constexpr int KB = Batch<T>::K;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    T rv[KB], pv[KB], qv[KB], ap[KB];
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      rv[kk] = V.r[g];
      ap[kk] = (T)A.p[0][g];
      if (MODE == 1 || !first) qv[kk] = V.q[g];
      if (MODE == 0 && !first) pv[kk] = V.p[g];
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      T v;
      if (MODE == 0) {
        if (first) {
          v = rv[kk];
        } else {
          const T tmp = pv[kk] + ng * qv[kk];  // AXPY(p, q_, -gamma)
          v = rv[kk] + beta * tmp;             // XPAY(p, r, beta)
        }
        V.p[g] = v;
      } else {
        v = na * qv[kk] + rv[kk];  // AXPYZ(s, q_, r, -alpha)
        V.s[g] = v;
      }
      const T m = omega * (v / ap[kk]);
      if (MODE == 0) V.pp[g] = m;
      else V.ss[g] = m;
      s[lds(tx, ty, k0 + kk)] = m;
    }
  }
  __syncthreads();
  T* h = MODE == 0 ? V.hpp : V.hss;
  push_halo<T, BI_NT>(B, cs[c], BC_NULL, slot, [&](int i, int j, int k) { return s[lds(i, j, k)]; }, h, h);
}

// P / S with a block red-black Gauss-Seidel preconditioner: v (p or s) as above in FP64, then
// z = one SOR sweep of M z = v from z = 0 in the matrix precision TM (FP32 copy of the matrix for
// the mixed-precision solve, FP64 for its all-FP64 counterpart), red cells first, the block's ghost values zero (a block-Jacobi / red-black GS
// preconditioner, fixed and linear); p_ (s_) = z widened to FP64, ghost layer pushed.
template <int MODE, class TM, class TZ = TM, bool SYM = false>
__global__ __launch_bounds__(BI_NT) void bi_precond_gs_kernel(const int* __restrict__ active,
                                                              const DevBlock* __restrict__ blk,
                                                              const DevCase* __restrict__ cs,
                                                              const SolveState* __restrict__ st, Coef7<TM> M,
                                                              BiVecs<double> V, int count) {
  // z of the block's own cells only: its ghost values are zero, so no ghost ring is staged and
  // every cell is written (red, then black) before a neighbour reads it
  __shared__ TZ z[NC];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const TZ omega = (TZ)st[c].omega;
  const double ng = -st[c].gamma, beta = st[c].beta, na = -st[c].alpha;
  // count 0: the iteration number comes from the case state (a replayed graph has no per-launch
  // arguments); FIN_BI_XR of the previous iteration stored it in iter
  const bool first = (count > 0 ? count : st[c].iter + 1) == 1 || st[c].restart;
  const int par = (tx + ty) & 1;  // red cells: k = 2m + par
  auto zi = [](int i, int j, int k) { return i + NB * (j + NB * k); };
  auto zat = [&](int i, int j, int k) -> TZ {
    return (i < 0 || i >= NB || j < 0 || j >= NB || k < 0 || k >= NB) ? (TZ)0 : z[zi(i, j, k)];
  };
  TZ vred[NB / 2], vblk[NB / 2];  // v of the red and of the black cells
  TM apr[NB / 2];
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += 4) {
    double rv[4], pv[4], qv[4];
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int k = k0 + kk;
      const size_t g = base + tx + NB * (ty + NB * k);
      rv[kk] = V.r[g];
      if (MODE == 1 || !first) qv[kk] = V.q[g];
      if (MODE == 0 && !first) pv[kk] = V.p[g];
      if ((k & 1) == par) apr[k >> 1] = M.p[0][g];
    }
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int k = k0 + kk;
      const size_t g = base + tx + NB * (ty + NB * k);
      double v;
      if (MODE == 0) {
        v = first ? rv[kk] : rv[kk] + beta * (pv[kk] + ng * qv[kk]);
        V.p[g] = v;
      } else {
        v = na * qv[kk] + rv[kk];
        V.s[g] = v;
      }
      if ((k & 1) == par) vred[k >> 1] = (TZ)v;
      else vblk[k >> 1] = (TZ)v;
    }
  }
#pragma unroll
  for (int m = 0; m < NB / 2; m++) {  // red: every neighbour is still zero
    const int k = 2 * m + par;
    z[zi(tx, ty, k)] = omega * (vred[m] / (TZ)apr[m]);
  }
  __syncthreads();
  // black (then, for the symmetric sweep, red again) from the updated neighbours
  auto half_sweep = [&](int colour_off) {
    const bool red = colour_off == par;
#pragma unroll
    for (int m0 = 0; m0 < NB / 2; m0 += 2) {
      TM a[2][7];
      TZ vv[2];
#pragma unroll
      for (int mm = 0; mm < 2; mm++) {
        const int k = 2 * (m0 + mm) + colour_off;
        const size_t g = base + tx + NB * (ty + NB * k);
#pragma unroll
        for (int q = 0; q < 7; q++) a[mm][q] = M.p[q][g];
        vv[mm] = red ? vred[m0 + mm] : vblk[m0 + mm];
      }
#pragma unroll
      for (int mm = 0; mm < 2; mm++) {
        const int k = 2 * (m0 + mm) + colour_off;
        const TZ rs = vv[mm] - (TZ)a[mm][1] * zat(tx - 1, ty, k) - (TZ)a[mm][2] * zat(tx + 1, ty, k) -
                      (TZ)a[mm][3] * zat(tx, ty - 1, k) - (TZ)a[mm][4] * zat(tx, ty + 1, k) -
                      (TZ)a[mm][5] * zat(tx, ty, k - 1) - (TZ)a[mm][6] * zat(tx, ty, k + 1);
        const TZ zo = z[zi(tx, ty, k)];
        z[zi(tx, ty, k)] = red ? zo + omega * (rs / (TZ)a[mm][0] - zo) : omega * (rs / (TZ)a[mm][0]);
      }
    }
    __syncthreads();
  };
  half_sweep(1 - par);
  if (SYM) half_sweep(par);
  double* out = MODE == 0 ? V.pp : V.ss;
#pragma unroll
  for (int k = 0; k < NB; k++) out[base + tx + NB * (ty + NB * k)] = (double)z[zi(tx, ty, k)];
  double* h = MODE == 0 ? V.hpp : V.hss;
  push_halo<double, BI_NT>(B, cs[c], BC_NULL, slot, [&](int i, int j, int k) { return (double)z[zi(i, j, k)]; }, h,
                           h);
}

// Compact matrix of the multi-sweep GS preconditioner, one 16-byte record per cell:
// x = 1 / Ap as FP32 bits, y / z / w = the normalised couplings A_q / Ap (q = 1..6: w e, s n, b t)
// as FP16 pairs (low half = odd q).  Any rounding here only defines a slightly different, still
// fixed and linear, preconditioner; the sweeps run in FP64.
__device__ inline uint32_t half_bits(double v) {
  const double lim = 65504.0;  // largest finite FP16
  const _Float16 h = (_Float16)(v > lim ? lim : v < -lim ? -lim : v);
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
__device__ inline double half_lo(uint32_t u) { return (double)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu)); }
__device__ inline double half_hi(uint32_t u) { return (double)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16)); }

__device__ inline uint4 gs_record(const Coef7<double>& A, size_t g) {
  const double d = A.p[0][g];
  const double inv = d != 0.0 ? 1.0 / d : 0.0;  // a zero pivot decouples the cell in M
  uint32_t w[6];
#pragma unroll
  for (int b = 0; b < 6; b++) w[b] = half_bits(A.p[b + 1][g] * inv);
  return make_uint4(__float_as_uint((float)inv), w[0] | (w[1] << 16), w[2] | (w[3] << 16), w[4] | (w[5] << 16));
}

__global__ __launch_bounds__(256) void gs_pack_kernel(const int* __restrict__ active, Coef7<double> A,
                                                      uint4* __restrict__ rec) {
  const size_t base = (size_t)active[blockIdx.x] * NC;
  for (int q = threadIdx.x; q < NC; q += 256) rec[base + q] = gs_record(A, base + q);
}

void launch_gs_pack(hipStream_t sm, const int* active, int nact, Coef7<double> A, uint4* rec) {
  if (nact) hipLaunchKernelGGL(gs_pack_kernel, dim3(nact), dim3(256), 0, sm, active, A, rec);
}

// Stencil position classes: a cell's class is (cx, cy, cz), each 0 on the block's low face layer,
// 2 on its high one and 1 inside (27 classes, index cx + 3 cy + 9 cz).  A block is class-uniform
// when every cell has bitwise the same seven coefficients as its class's representative cell
// (the lowest-index cell of the class); a block without cut cells is, whatever its outer faces
// and level interfaces (those change the coefficients of a whole face layer alike).  For such a
// block the 27 rows -- as FP32 bands (8 floats each, the eighth unused) and as preconditioner
// records -- stand for the per-cell arrays, bit for bit.
__device__ inline int stencil_class1(int i) { return i == 0 ? 0 : i == NB - 1 ? 2 : 1; }
__device__ inline int stencil_rep1(int c) { return c == 0 ? 0 : c == 2 ? NB - 1 : 1; }

__global__ __launch_bounds__(256) void stencil_class_kernel(const int* __restrict__ active, Coef7<double> A,
                                                            int* __restrict__ uniform, float* __restrict__ tab,
                                                            uint4* __restrict__ trec) {
  const int slot = active[blockIdx.x];
  const size_t base = (size_t)slot * NC;
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const int rx = stencil_rep1(stencil_class1(tx)), ry = stencil_rep1(stencil_class1(ty));
  bool same = true;
  for (int k = 0; k < NB; k++) {
    const size_t g = base + tx + NB * (ty + NB * k);
    const size_t r = base + rx + NB * (ry + NB * stencil_rep1(stencil_class1(k)));
#pragma unroll
    for (int q = 0; q < 7; q++)
      same = same && __double_as_longlong(A.p[q][g]) == __double_as_longlong(A.p[q][r]);
  }
  const int all = __syncthreads_and(same ? 1 : 0);
  if (t < 27) {
    const size_t r = base + stencil_rep1(t % 3) + NB * (stencil_rep1((t / 3) % 3) + NB * stencil_rep1(t / 9));
    float* row = tab + ((size_t)slot * 27 + t) * 8;
#pragma unroll
    for (int q = 0; q < 7; q++) row[q] = (float)A.p[q][r];
    row[7] = 0.0f;
    trec[(size_t)slot * 27 + t] = gs_record(A, r);
  }
  if (t == 0) uniform[slot] = all;
}

void launch_stencil_classes(hipStream_t sm, const int* active, int nact, Coef7<double> A, int* uniform, float* tab,
                            uint4* trec) {
  if (nact) hipLaunchKernelGGL(stencil_class_kernel, dim3(nact), dim3(256), 0, sm, active, A, uniform, tab, trec);
}

// P / S with NS red-black SOR sweeps of the block-local system (ghost values zero) from z = 0,
// on the compact matrix above, in FP64: u = v / Ap; red cells first; each update
// z <- z + omega (u - sum_q w_q z_q - z).  The block's z lives in LDS; each of the GSK_NT threads
// keeps u and the couplings of half an (i, j) column (8 cells) in registers, so the sweeps after
// the first read no HBM.  The couplings to cells outside the block are zeroed in the registers
// (ghost values are zero in the block-local system), so a sweep reads its six LDS neighbours at
// fixed offsets with no bounds test: the x / y offsets of an edge cell then wrap to a cell of
// the same colour of another column (multiplied by zero) and the k offsets of the end planes to
// a zero plane below and above the block.
constexpr int GSK_NT = 512;
constexpr int GSK_NK = NB * 256 / GSK_NT;  // cells of a column per thread
// vs_out (optional): the block sum of v (restriction P^T v of the coarse correction).
// MUL (multiplicative two-level): e (this mode's coarse correction per slot, computed before this
// kernel) is subtracted first, the sweeps smooth r1 = v - A P e with A the compact matrix
// (A_0 = 1 / dinv, A_q = w_q A_0): u1 = u - (e_a + sum_q w_q e of the block each neighbour
// lies in).
// REG 1 (every block of the launch class-uniform): the records are the block's class records
// trec[slot][27] instead of the per-cell array, and the sweeps keep one set of couplings for the
// thread's inner cells and one for its end cell (k = 0 or NB - 1) instead of one per cell, which
// leaves registers for three workgroups per CU.  REG 2 (mixed launch): each block takes its
// records from the class rows when uni[slot] is set, else per cell.
template <class T, int MODE, int NS, bool MUL, int REG>
__global__ __launch_bounds__(GSK_NT) __attribute__((amdgpu_waves_per_eu(REG == 1 ? 6 : 1))) void bi_precond_gsk_kernel(const int* __restrict__ active,
                                                               const DevBlock* __restrict__ blk,
                                                               const DevCase* __restrict__ cs,
                                                               const SolveState* __restrict__ st,
                                                               const uint4* __restrict__ rec, BiVecs<T> V,
                                                               int count, double* __restrict__ vs_out,
                                                               const double* __restrict__ ce,
                                                               const uint4* __restrict__ trec,
                                                               const int* __restrict__ uni) {
  __shared__ double zpad[NC + 2 * NF];  // z with a zero plane below (k = -1) and above (k = NB)
  __shared__ double red[GSK_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int c = B.case_id;
  if (st[c].done) return;
  double* const z = zpad + NF;
  const int t = threadIdx.x, tx = t & 15, ty = (t >> 4) & 15;
  const int kb = (t >> 8) * GSK_NK;  // this thread's cells: k = kb .. kb + GSK_NK - 1
  static_assert(GSK_NT == 2 * NF, "one pad plane entry per thread");
  zpad[t < NF ? t : NC + t] = 0.0;
  const size_t base = (size_t)slot * NC;
  // the block's vectors from uniform base pointers, indexed by 32-bit cell offsets (one offset
  // register serves every vector)
  const unsigned o0 = tx + NB * ty + NF * kb;
  const T* const rb = V.r + base;
  const T* const qb = V.q + base;
  T* const vb = (MODE == 0 ? V.p : V.s) + base;  // MODE 0 reads the old p here too
  const uint4* const recb = rec + base;
  const double omega = st[c].omega;
  const double ng = -st[c].gamma, beta = st[c].beta, na = -st[c].alpha;
  // count 0: the iteration number comes from the case state (a replayed graph has no per-launch
  // arguments); FIN_BI_XR of the previous iteration stored it in iter
  const bool first = (count > 0 ? count : st[c].iter + 1) == 1 || st[c].restart;
  const int par = (tx + ty) & 1;  // red cells of this column: k = kb + 2m + par (kb even)
  auto zi = [](int i, int j, int k) { return i + NB * (j + NB * k); };
  // per colour (index m: k = kb + 2m + par for red, kb + 2m + 1 - par for black), compile-time
  // indexed so that they stay in registers
  double ur[GSK_NK / 2], ub[GSK_NK / 2];
  uint32_t wr[GSK_NK / 2][3], wb[GSK_NK / 2][3];
  double vsum = 0.0;
  // couplings kept for the sweeps: those to cells outside the block zeroed (per half word)
  const uint32_t mx = (tx == 0 ? 0xffff0000u : ~0u) & (tx == NB - 1 ? 0x0000ffffu : ~0u);
  const uint32_t my = (ty == 0 ? 0xffff0000u : ~0u) & (ty == NB - 1 ? 0x0000ffffu : ~0u);
  // REG: the class records of the thread's inner cells and of its end cell (k = 0 for the lower
  // half column, NB - 1 for the upper), and their couplings as kept for the sweeps
  const bool cls = REG == 1 || (REG == 2 && uni[slot] != 0);  // uniform over the workgroup
  uint4 rin = {}, rend = {};
  uint32_t win[3] = {}, wend[3] = {};
  if (REG != 0 && cls) {
    const uint4* tr = trec + (size_t)slot * 27 + stencil_class1(tx) + 3 * stencil_class1(ty);
    rin = tr[9];
    rend = tr[kb == 0 ? 0 : 18];
    win[0] = rin.y & mx;
    win[1] = rin.z & my;
    win[2] = rin.w;
    wend[0] = rend.y & mx;
    wend[1] = rend.z & my;
    wend[2] = rend.w & (kb == 0 ? 0xffff0000u : 0x0000ffffu);
  }
  const bool lower = kb == 0;
  // MUL: e of this block (eo) and of the block the ghost value of the x, y and z face this
  // thread's cells touch comes from (ex: face x- when tx = 0, else x+; ey likewise; ez: z- for the
  // lower half column, z+ for the upper).  A half column never straddles a subface boundary, so
  // one value per axis covers the thread's 8 cells; the indices are formed for every lane (a lane
  // off the face does not use its value) so that the loads issue together.
  double eo = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
  if constexpr (MUL) {
    const int sx = ghost_source(B, slot, tx == 0 ? 0 : 1, ty, kb);
    const int sy = ghost_source(B, slot, ty == 0 ? 2 : 3, tx, kb);
    const int sz = ghost_source(B, slot, kb == 0 ? 4 : 5, tx, ty);
    eo = ce[slot];
    ex = ce[sx];
    ey = ce[sy];
    ez = ce[sz];
  }
#pragma unroll
  for (int k0 = 0; k0 < GSK_NK; k0 += 4) {
    T rv[4], pv[4], qv[4];
    uint4 rc[4];
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const unsigned o = o0 + NF * (k0 + kk);
      // streamed once through this kernel: non-temporal loads and stores leave the caches to what
      // the next kernels read (measured: this kernel 3% faster, the matvecs after it 3-7%)
      rv[kk] = __builtin_nontemporal_load(rb + o);
      if (MODE == 1 || !first) qv[kk] = __builtin_nontemporal_load(qb + o);
      if (MODE == 0 && !first) pv[kk] = __builtin_nontemporal_load(vb + o);
      if (REG == 0 || (REG == 2 && !cls)) rc[kk] = recb[o];
    }
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int k = k0 + kk;  // relative to kb
      if (REG != 0 && cls) {  // the end cell of the half column is k = 0 (lower) or GSK_NK - 1 (upper)
        if (k == 0) rc[kk] = lower ? rend : rin;
        else if (k == GSK_NK - 1) rc[kk] = lower ? rin : rend;
        else rc[kk] = rin;
      }
      // FP64 vectors: the recurrence's own values; FP32 vectors (an inner solve of a defect
      // correction): formed in FP64 and rounded once when stored
      double v;
      if (MODE == 0)
        v = first ? (double)rv[kk] : (double)rv[kk] + beta * ((double)pv[kk] + ng * (double)qv[kk]);
      else
        v = na * (double)qv[kk] + (double)rv[kk];
      if constexpr (sizeof(T) == 4) v = (double)(T)v;  // the value the recurrence keeps
      __builtin_nontemporal_store((T)v, vb + o0 + NF * k);
      vsum += v;
      double uk = v * (double)__uint_as_float(rc[kk].x);
      const int ka = kb + k;
      if constexpr (MUL) {
        uk -= eo + half_lo(rc[kk].y) * (tx > 0 ? eo : ex) + half_hi(rc[kk].y) * (tx < NB - 1 ? eo : ex) +
              half_lo(rc[kk].z) * (ty > 0 ? eo : ey) + half_hi(rc[kk].z) * (ty < NB - 1 ? eo : ey) +
              half_lo(rc[kk].w) * (ka > 0 ? eo : ez) + half_hi(rc[kk].w) * (ka < NB - 1 ? eo : ez);
      }
      const int m = k >> 1;
      const uint32_t mz = (ka == 0 ? 0xffff0000u : ~0u) & (ka == NB - 1 ? 0x0000ffffu : ~0u);
      if ((k & 1) == par) {
        ur[m] = uk;
        if constexpr (REG != 1) {
          wr[m][0] = rc[kk].y & mx;
          wr[m][1] = rc[kk].z & my;
          wr[m][2] = rc[kk].w & mz;
        }
      } else {
        ub[m] = uk;
        if constexpr (REG != 1) {
          wb[m][0] = rc[kk].y & mx;
          wb[m][1] = rc[kk].z & my;
          wb[m][2] = rc[kk].w & mz;
        }
      }
    }
  }
  if (vs_out) {
    const double sv = block_sum<GSK_NT>(vsum, red);
    if (t == 0) vs_out[slot] = sv;
  }
  // first red half sweep: every neighbour is still zero; the black cells start at zero (the
  // wrapped reads of the first black half sweep must see finite values)
#pragma unroll
  for (int m = 0; m < GSK_NK / 2; m++) {
    z[zi(tx, ty, kb + 2 * m + par)] = omega * ur[m];
    z[zi(tx, ty, kb + 2 * m + 1 - par)] = 0.0;
  }
  __syncthreads();
  // one half sweep over this column's cells of one colour; from_zero: the cells' old values are
  // zero (first visit)
  auto half = [&](auto red_tag, bool from_zero) {
    constexpr bool RED = decltype(red_tag)::value;
    const int cp = RED ? par : 1 - par;  // this colour's cells: k = kb + 2m + cp
    const int kp = kb + cp;
#pragma unroll
    for (int m = 0; m < GSK_NK / 2; m++) {
      const int k = 2 * m + kp;
      uint32_t w[3];
      if constexpr (REG == 1) {
        // the end cell: m = 0 with cp = 0 (lower half), m = last with cp = 1 (upper half)
        const bool end = m == 0 ? (lower && cp == 0) : m == GSK_NK / 2 - 1 ? (!lower && cp == 1) : false;
#pragma unroll
        for (int b = 0; b < 3; b++) w[b] = end ? wend[b] : win[b];
      } else {
#pragma unroll
        for (int b = 0; b < 3; b++) w[b] = RED ? wr[m][b] : wb[m][b];
      }
      // keep the couplings packed across sweeps (an empty asm hides that the unpacked values
      // of the last sweep could be reused, which would cost 12 more registers per cell)
      asm volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]));
      const double* zc = z + zi(tx, ty, k);
      double sum = half_lo(w[0]) * zc[-1];
      sum = __builtin_fma(half_hi(w[0]), zc[1], sum);
      sum = __builtin_fma(half_lo(w[1]), zc[-NB], sum);
      sum = __builtin_fma(half_hi(w[1]), zc[NB], sum);
      sum = __builtin_fma(half_lo(w[2]), zc[-NB * NB], sum);
      sum = __builtin_fma(half_hi(w[2]), zc[NB * NB], sum);
      const double zn = (RED ? ur[m] : ub[m]) - sum;
      if (from_zero) {
        z[zi(tx, ty, k)] = omega * zn;
      } else {
        const double zo = z[zi(tx, ty, k)];
        z[zi(tx, ty, k)] = __builtin_fma(omega, zn - zo, zo);
      }
    }
    __syncthreads();
  };
  using Red = std::true_type;
  using Black = std::false_type;
  half(Black{}, true);
#pragma unroll
  for (int sweep = 1; sweep < NS; sweep++) {
    half(Red{}, false);
    half(Black{}, false);
  }
  // MUL: the stored direction is z + P e (e added to this block's values and to the ghost values
  // it pushes, which then come from z + e through every face kind)
  T* const out = (MODE == 0 ? V.pp : V.ss) + base;
  unsigned oo = o0;  // opaque: the store offsets are formed here, not held across the sweeps
  asm volatile("" : "+v"(oo));
#pragma unroll
  for (int k = 0; k < GSK_NK; k++)
    __builtin_nontemporal_store((T)(MUL ? z[oo + NF * k] + eo : z[oo + NF * k]), out + oo + NF * k);
  T* h = MODE == 0 ? V.hpp : V.hss;
  push_halo<T, GSK_NT>(
      B, cs[c], BC_NULL, slot, [&](int i, int j, int k) { return MUL ? z[zi(i, j, k)] + eo : z[zi(i, j, k)]; }, h,
      h);
}

// Launches the sweep kernel; with the class rows (o.trec) the class-uniform variant when every
// active block is class-uniform, the mixed one when some are, else the per-cell one.
template <class T, int MODE, bool MUL>
void launch_gsk(const BiOpts& o, hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                const SolveState* st, BiVecs<T> V, int count, double* vs_out, const double* ce) {
  const dim3 g(nact), b(GSK_NT);
  const int reg = !o.trec || o.nreg == 0 ? 0 : o.nirr == 0 ? 1 : 2;
#define GSK1(NS, R)                                                                                                 \
  hipLaunchKernelGGL((bi_precond_gsk_kernel<T, MODE, NS, MUL, R>), g, b, 0, sm, active, blk, cs, st, o.rec, V, count, \
                     vs_out, ce, o.trec, o.uni)
#define GSK(NS)            \
  do {                     \
    if (reg == 1)          \
      GSK1(NS, 1);         \
    else if (reg == 2)     \
      GSK1(NS, 2);         \
    else                   \
      GSK1(NS, 0);         \
  } while (0)
  switch (o.sweeps) {
    case 1: GSK(1); break;
    case 2: GSK(2); break;
    case 3: GSK(3); break;
    case 4: GSK(4); break;
    case 5: GSK(5); break;
    default: GSK(6);
  }
#undef GSK
#undef GSK1
}

// Q: q_ = A p_, q_.r0 -> alpha (mode 0);  T: t_ = A s_, t_.s and t_.t_ -> gamma (mode 1)
// CC: the operand's values (interior and ghosts) get the coarse correction ce[MODE * nslots + the
// source block] added (two-level preconditioners).
// LDS holds the operand's interior (NC values) and the ghost layers of the four side faces
// (4 NF values, same layout as the halo); the two z ghosts of a column stay in registers.  In
// FP64 that is 40 KiB, so four workgroups fit a CU.
// REG 1 (FP32 bands, every block of the launch class-uniform): the bands come from the block's 27
// class rows tab[slot][27][8] (the thread's column needs three: k = 0, inner, NB - 1), so no
// per-cell band is read; REG 2 (mixed launch): so for the blocks with uni[slot] set.
template <class T, int MODE, class TA = T, bool HR = false, bool CC = false, int REG = 0>
__global__ __launch_bounds__(BI_NT) void bi_matvec_kernel(const int* __restrict__ active,
                                                          const DevBlock* __restrict__ blk,
                                                          const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                          Coef7<TA> A, BiVecs<T> V, double* __restrict__ part,
                                                          size_t nslots, const double* __restrict__ ce = nullptr,
                                                          double* __restrict__ bsy = nullptr,
                                                          const float* __restrict__ tab = nullptr,
                                                          const int* __restrict__ uni = nullptr, int early = 0) {
  __shared__ T s[NC + 4 * NF];
  double* const red = reinterpret_cast<double*>(s);  // after the last read of s (block_sum syncs first)
  static_assert(NC * sizeof(T) >= BI_NT / 64 * sizeof(double), "reduction scratch fits");
  static_assert(BI_NT == NF, "one thread per (i, j) column and per face position");
  const int slot = active[blockIdx.x];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  const T* __restrict__ w = MODE == 0 ? V.r0 : V.s;
  T* __restrict__ y = MODE == 0 ? V.q : V.t;
  const int o0 = tx + NB * ty;
  constexpr int KB = Batch<T>::K;
  double d0 = 0.0, d1 = 0.0, ys = 0.0;
  // The operand's values and ghosts (and the class rows) depend on the slot alone.  early (a launch
  // of one case, whose blocks finish together): they are loaded before the case state is read, so
  // the chain of dependent reads (block -> case -> state) overlaps the loads instead of preceding
  // them (every address is inside its own block; the empty asm with a memory clobber keeps the
  // compiler from sinking the loads past the test and waits for nothing).  In a launch of several
  // cases, which finish at different iterations, a finished case's workgroups would load for
  // nothing: there they load after the test.
  const T* __restrict__ xo = (MODE == 0 ? V.pp : V.ss) + base;
  const T* __restrict__ h = (MODE == 0 ? V.hpp : V.hss) + (size_t)slot * NH;
  T v[NB], gv[4], zm, zp;  // zm / zp: the ghosts below k = 0 / above k = NB - 1
  // REG: the bands of the column's classes (k = 0, inner, NB - 1)
  const bool cls = REG == 1 || (REG == 2 && uni[slot] != 0);  // uniform over the workgroup
  TA al[7], ai[7], ah[7];
  auto stage_loads = [&]() {
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = xo[o0 + NF * k];
#pragma unroll
    for (int f = 0; f < 4; f++) gv[f] = h[f * NF + t];
    zm = h[halo_index(4, tx, ty)];
    zp = h[halo_index(5, tx, ty)];
    if (REG != 0 && cls) {
      static_assert(REG == 0 || sizeof(TA) == 4, "class rows are FP32 bands");
      const float* tr = tab + ((size_t)slot * 27 + stencil_class1(tx) + 3 * stencil_class1(ty)) * 8;
#pragma unroll
      for (int q = 0; q < 7; q++) {
        al[q] = tr[q];
        ai[q] = tr[9 * 8 + q];
        ah[q] = tr[18 * 8 + q];
      }
    }
  };
  if (early) {
    stage_loads();
    asm volatile("" ::: "memory");
  }
  const int c = blk[slot].case_id;
  if (st[c].done) return;
  if (!early) stage_loads();
  const uint32_t local = (uint32_t)(slot - cs[c].slot_begin);
  T a[KB][7], wv[KB];
  auto load_batch = [&](int k0) {
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const int k = k0 + kk;
      const size_t g = base + o0 + NF * k;
      if (REG != 0 && cls) {
#pragma unroll
        for (int q = 0; q < 7; q++) a[kk][q] = (T)(k == 0 ? al[q] : k == NB - 1 ? ah[q] : ai[q]);
      } else {
#pragma unroll
        for (int q = 0; q < 7; q++) a[kk][q] = (T)A.p[q][g];
      }
      wv[kk] = (MODE == 0 && HR) ? (T)shadow_value(local, (uint32_t)(g - base)) : w[g];
    }
  };
  load_batch(0);  // in flight while the operand is staged
  if constexpr (CC) {  // each thread corrects the values it stages
    const DevBlock& Bk = blk[slot];
    const double* e = ce + MODE * nslots;
    const T eo = (T)e[slot];
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] += eo;
#pragma unroll
    for (int f = 0; f < 4; f++) gv[f] += (T)e[ghost_source(Bk, slot, f, t & 15, t >> 4)];
    zm += (T)e[ghost_source(Bk, slot, 4, tx, ty)];
    zp += (T)e[ghost_source(Bk, slot, 5, tx, ty)];
  }
#pragma unroll
  for (int k = 0; k < NB; k++) s[o0 + NF * k] = v[k];
#pragma unroll
  for (int f = 0; f < 4; f++) s[NC + f * NF + t] = gv[f];
  __syncthreads();
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    if (k0 > 0) load_batch(k0);
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const int k = k0 + kk;
      const int li = o0 + NF * k;
      const size_t g = base + li;
      // a face lane reads the ghost layer (halo layout: face f at (a, b) = NC + f NF + a + NB b)
      const T vxm = s[tx > 0 ? li - 1 : NC + halo_index(0, ty, k)];
      const T vxp = s[tx < NB - 1 ? li + 1 : NC + halo_index(1, ty, k)];
      const T vym = s[ty > 0 ? li - NB : NC + halo_index(2, tx, k)];
      const T vyp = s[ty < NB - 1 ? li + NB : NC + halo_index(3, tx, k)];
      const T vzm = k > 0 ? s[li - NF] : zm;
      const T vzp = k < NB - 1 ? s[li + NF] : zp;
      // calc_ax order (bils.f90:195-202)
      const T r = a[kk][1] * vxm + a[kk][2] * vxp + a[kk][3] * vym + a[kk][4] * vyp + a[kk][5] * vzm +
                  a[kk][6] * vzp + a[kk][0] * s[li];
      y[g] = r;
      ys += (double)r;
      d0 += (double)r * (double)wv[kk];
      if (MODE == 1) d1 += (double)r * (double)r;
    }
  }
  double s0 = block_sum<BI_NT>(d0, red);
  double s1 = MODE == 1 ? block_sum<BI_NT>(d1, red) : 0.0;
  if (t == 0) {
    part[slot] = s0;
    part[nslots + slot] = s1;
  }
  if (bsy) {  // block sums of q_ / t_ (multiplicative coarse correction)
    const double v = block_sum<BI_NT>(ys, red);
    if (t == 0) bsy[slot] = v;
  }
}

// CC: p_ and s_ carry the coarse corrections ce[slot] / ce[nslots + slot] (two-level
// preconditioners), added as the matvecs added them.
// tpart: the T stage's block partials (t_.s, t_.t_); every workgroup adds its case's partials in
// the fixed order of the per-case finalisation and takes gamma = (t_.s) / (t_.t_)
// (FFVILS.h:905-915) itself, the case's first block also storing it in the case state, so no
// finalisation launch sits between T and XR.  The case's first block also copies this
// iteration's rho (rr1) to rr0, as the iteration's finalisation does, so that a finalisation formed
// inside the next kernel can read it there while rr1 is rewritten.
template <class T, bool HR = false, bool CC = false>
__global__ __launch_bounds__(BI_NT) void bi_xr_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                      const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                      T* __restrict__ x, BiVecs<T> V, double* __restrict__ part,
                                                      size_t nslots, int count, const double* __restrict__ ce = nullptr,
                                                      double* __restrict__ bsr = nullptr,
                                                      const double* __restrict__ tpart = nullptr, int early = 0) {
  __shared__ double red[2 * (BI_NT / 64)];
  __shared__ double gshared;
  const int slot = active[blockIdx.x];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const size_t base = (size_t)slot * NC;
  constexpr int KB = Batch<T>::K;
  T pv[KB], sv_[KB], xv[KB], tv[KB], sv[KB], r0v[KB];
  auto load_batch = [&](int k0) {
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      // the last reads of p_, s_, t_ and s in the iteration, and x's only one: non-temporal
      // (measured 3% faster; r0, read every iteration, stays cached)
      pv[kk] = __builtin_nontemporal_load(V.pp + g);
      sv_[kk] = __builtin_nontemporal_load(V.ss + g);
      xv[kk] = __builtin_nontemporal_load(x + g);
      tv[kk] = __builtin_nontemporal_load(V.t + g);
      sv[kk] = __builtin_nontemporal_load(V.s + g);
      if constexpr (!HR) r0v[kk] = V.r0[g];
    }
  };
  // The first batch depends on the slot alone.  early (a launch of one case, whose blocks finish
  // together): it is issued before the case state is read, so the chain of dependent reads (block
  // -> case -> state) overlaps the loads instead of preceding them (every address is inside its own
  // block; the empty asm with a memory clobber keeps the compiler from sinking the loads past the
  // test and waits for nothing).  In a launch of several cases, which finish at different
  // iterations, a finished case's workgroups would load for nothing: there it follows the test.
  if (early) {
    load_batch(0);
    asm volatile("" ::: "memory");
  }
  const int c = blk[slot].case_id;
  if (st[c].done) return;
  if (!early) load_batch(0);
  // gamma: formed after the first batch of loads is issued (the partials' round trips then
  // overlap the block's own loads)
  auto form_gamma = [&]() -> double {
    if (tpart) {
      const DevCase& C = cs[c];
      double t0, t1;
      case_sum2<BI_NT>(tpart, tpart + nslots, C.slot_begin, C.nblk, red, t0, t1);
      if (t == 0) {
        gshared = t0 / t1;
        if (slot == C.slot_begin) st[c].gamma = gshared;
      }
      __syncthreads();
      return gshared;
    }
    return st[c].gamma;
  };
  T alpha = 0, gamma = 0, ng = 0;
  const int sb = cs[c].slot_begin;
  const uint32_t local = (uint32_t)(slot - sb);
  double rr = 0.0, rr1 = 0.0, rs = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < NB; k0 += KB) {
    if (k0 > 0) load_batch(k0);
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      if constexpr (CC) {
        pv[kk] += (T)ce[slot];
        sv_[kk] += (T)ce[nslots + slot];
      }
      if constexpr (HR) r0v[kk] = (T)shadow_value(local, (uint32_t)(tx + NB * (ty + NB * (k0 + kk))));
    }
    if (k0 == 0) {
      if (t == 0 && slot == sb) st[c].rr0 = st[c].rr1;
      const double gd = form_gamma();
      alpha = (T)st[c].alpha;
      gamma = (T)gd;
      ng = (T)(-gd);
    }
#pragma unroll
    for (int kk = 0; kk < KB; kk++) {
      const size_t g = base + tx + NB * (ty + NB * (k0 + kk));
      __builtin_nontemporal_store(alpha * pv[kk] + gamma * sv_[kk] + xv[kk], x + g);  // AXPBYPZ(x, p_, s_, alpha, gamma)
      const T r = ng * tv[kk] + sv[kk];                  // AXPYZ(r, t_, s, -gamma)
      V.r[g] = r;
      rs += (double)r;
      rr += (double)r * (double)r;
      rr1 += (double)r * (double)r0v[kk];
    }
  }
  double s0 = block_sum<BI_NT>(rr, red);
  double s1 = block_sum<BI_NT>(rr1, red);
  if (t == 0) {
    part[slot] = s0;
    part[nslots + slot] = s1;
  }
  if (bsr) {
    const double v = block_sum<BI_NT>(rs, red);
    if (t == 0) bsr[slot] = v;
  }
}

// Reliable update (mixed precision): r has just been replaced by the true residual of the FP64
// solution; r.r becomes the new reference norm and r.r0 the new rho (part[0], part[1]).
template <class T>
__global__ __launch_bounds__(BI_NT) void bi_reanchor_kernel(const int* __restrict__ active,
                                                            const DevBlock* __restrict__ blk,
                                                            const SolveState* __restrict__ gate, BiVecs<T> V,
                                                            double* __restrict__ part, size_t nslots) {
  __shared__ double red[BI_NT / 64];
  const int slot = active[blockIdx.x];
  if (gate[blk[slot].case_id].done) return;
  const size_t base = (size_t)slot * NC;
  double rr = 0.0, rr0 = 0.0;
  for (int q = threadIdx.x; q < NC; q += BI_NT) {
    const double r = (double)V.r[base + q];
    rr += r * r;
    rr0 += r * (double)V.r0[base + q];
  }
  const double s0 = block_sum<BI_NT>(rr, red), s1 = block_sum<BI_NT>(rr0, red);
  if (threadIdx.x == 0) {
    part[slot] = s0;
    part[nslots + slot] = s1;
  }
}

template <class T>
__global__ __launch_bounds__(BI_NT) void bi_shadow_kernel(const int* __restrict__ active,
                                                          const DevBlock* __restrict__ blk,
                                                          const DevCase* __restrict__ cs,
                                                          const SolveState* __restrict__ gate, BiVecs<T> V) {
  const int slot = active[blockIdx.x];
  const int c = blk[slot].case_id;
  if (gate[c].done) return;
  const uint32_t local = (uint32_t)(slot - cs[c].slot_begin);
  const size_t base = (size_t)slot * NC;
  for (int q = threadIdx.x; q < NC; q += BI_NT) V.r0[base + q] = (T)shadow_value(local, (uint32_t)q);
}

template <class T>
void launch_bicgstab_shadow(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                            const SolveState* gate, BiVecs<T> V) {
  if (nact) hipLaunchKernelGGL(bi_shadow_kernel<T>, dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, gate, V);
}

// Restart after a breakdown at a reliable update: r0 := r for the cases that asked for it.
template <class T>
__global__ __launch_bounds__(BI_NT) void bi_restart_kernel(const int* __restrict__ active,
                                                           const DevBlock* __restrict__ blk,
                                                           const SolveState* __restrict__ st, BiVecs<T> V) {
  const int slot = active[blockIdx.x];
  if (!st[blk[slot].case_id].restart) return;
  const size_t base = (size_t)slot * NC;
  for (int q = threadIdx.x; q < NC; q += BI_NT) V.r0[base + q] = V.r[base + q];
}

template <class T>
void launch_bicgstab_reanchor(hipStream_t sm, const int* active, int nact, const DevBlock* blk,
                              const SolveState* gate, BiVecs<T> V, double* part, size_t nslots) {
  if (nact) hipLaunchKernelGGL(bi_reanchor_kernel<T>, dim3(nact), dim3(BI_NT), 0, sm, active, blk, gate, V, part, nslots);
}

template <class T>
void launch_bicgstab_restart(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const SolveState* st,
                             BiVecs<T> V) {
  if (nact) hipLaunchKernelGGL(bi_restart_kernel<T>, dim3(nact), dim3(BI_NT), 0, sm, active, blk, st, V);
}

template <class T>
void launch_bicgstab_init(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                          SolveState* st, Coef7<T> A, const T* b, const T* x, const T* hx, BiVecs<T> V, double* part,
                          size_t nslots, bool hash_r0, double* bsr) {
  if (!nact) return;
  if (hash_r0)
    hipLaunchKernelGGL((bi_init_kernel<T, true>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, A, b, x, hx, V,
                       part, nslots, bsr);
  else
    hipLaunchKernelGGL((bi_init_kernel<T, false>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, A, b, x, hx,
                       V, part, nslots, bsr);
}

template <class T>
void launch_bicgstab_iter(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase,
                          const DevBlock* blk, const DevCase* cs, SolveState* st, Coef7<T> A, T* x, BiVecs<T> V,
                          double* part, size_t nslots, int count, int mask, const BiOpts& o) {
  // mask bits 0-4 select P, Q, S, T, XR; bit 5 leaves out the per-case finalisations; bit 8
  // splits the coarse corrections of the two-level methods into stages of their own (bits 6, 7).
  // Multiplicative two-level methods (count > 0): bit 9 has this iteration's first coarse
  // application form the previous iteration's finalisation from the XR partials still in part[0..1]
  // (whose launch bit 10 left out).
  if (!nact) return;
  const bool dbl = sizeof(T) == 8;
  // FP32 vectors (the inner solve of a defect correction) take the compact-record path only with
  // the multiplicative coarse correction (methods 23-28's preconditioner)
  const bool gsc_mult = o.gs == BiOpts::GSC && o.cinv && o.mult;
  if (!(gsc_mult && count > 0) || (mask & 256)) mask &= ~(512 | 1024);
  // XR and the matvecs load before reading the case state only in a launch of one case
  const int early = ncase == 1 ? 1 : 0;
  BiVecs<double>& Vd = *reinterpret_cast<BiVecs<double>*>(&V);
  double* const part0 = part;
  auto precond = [&](auto mode_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    if (gsc_mult)  // block sum of p kept for the next coarse rhs
      launch_gsk<T, MODE, true>(o, sm, active, nact, blk, cs, st, V, count, MODE == 0 ? o.ce + 4 * nslots : nullptr,
                                o.ce + MODE * nslots);
    else if (dbl && o.gs == BiOpts::GSC && o.cinv)
      launch_gsk<double, MODE, false>(o, sm, active, nact, blk, cs, st, Vd, count, part + 4 * nslots, nullptr);
    else if (dbl && o.gs == BiOpts::GSC)
      launch_gsk<double, MODE, false>(o, sm, active, nact, blk, cs, st, Vd, count, nullptr, nullptr);
    else if (dbl && o.gs == BiOpts::GS32)
      hipLaunchKernelGGL((bi_precond_gs_kernel<MODE, float>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st,
                         *o.M32, Vd, count);
    else if (dbl && o.gs == BiOpts::GS32_ARITH64)
      hipLaunchKernelGGL((bi_precond_gs_kernel<MODE, float, double>), dim3(nact), dim3(BI_NT), 0, sm, active, blk,
                         cs, st, *o.M32, Vd, count);
    else if (dbl && o.gs == BiOpts::SGS32_ARITH64)
      hipLaunchKernelGGL((bi_precond_gs_kernel<MODE, float, double, true>), dim3(nact), dim3(BI_NT), 0, sm, active,
                         blk, cs, st, *o.M32, Vd, count);
    else if (dbl && o.gs == BiOpts::GS64)
      hipLaunchKernelGGL((bi_precond_gs_kernel<MODE, double>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st,
                         *o.M64, Vd, count);
    else if (dbl && o.A32)
      hipLaunchKernelGGL((bi_precond_kernel<double, MODE, float>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs,
                         st, *o.A32, Vd, count);
    else
      hipLaunchKernelGGL((bi_precond_kernel<T, MODE>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, A, V,
                         count);
  };
  // the T stage keeps its partials in part[2..3] (XR reads them while writing part[0..1])
  auto matvec = [&](auto mode_tag) {
    constexpr int MODE = decltype(mode_tag)::value;
    double* const part = MODE == 0 ? part0 : part0 + 2 * nslots;
    if ((dbl || gsc_mult) && o.A32 && o.gs == BiOpts::GSC) {
      // 11-28: the additive correction (17-22) adds e to the operand; the multiplicative one
      // (23-28) finds it in p_ / s_ already and keeps the block sum of q_.  Class-uniform blocks
      // take their bands from the class rows.
      double* const bsy = (o.cinv && o.mult && MODE == 0) ? o.ce + 3 * nslots : nullptr;
      const int reg = !o.trec || o.nreg == 0 ? 0 : o.nirr == 0 ? 1 : 2;
      auto go = [&](auto hr_tag, auto cc_tag, auto reg_tag) {
        constexpr bool HR = decltype(hr_tag)::value, CC = decltype(cc_tag)::value;
        constexpr int REG = decltype(reg_tag)::value;
        hipLaunchKernelGGL((bi_matvec_kernel<T, MODE, float, HR, CC, REG>), dim3(nact), dim3(BI_NT), 0, sm,
                           active, blk, cs, st, *o.A32, V, part, nslots, CC ? (const double*)o.ce : nullptr, bsy,
                           o.tab, o.uni, early);
      };
      auto regs = [&](auto hr_tag, auto cc_tag) {
        if (reg == 1)
          go(hr_tag, cc_tag, std::integral_constant<int, 1>{});
        else if (reg == 2)
          go(hr_tag, cc_tag, std::integral_constant<int, 2>{});
        else
          go(hr_tag, cc_tag, std::integral_constant<int, 0>{});
      };
      auto ccs = [&](auto hr_tag) {
        if constexpr (sizeof(T) == 8) {
          if (o.cinv && !o.mult) {
            regs(hr_tag, std::true_type{});
            return;
          }
        }
        regs(hr_tag, std::false_type{});
      };
      if (o.hash_r0)
        ccs(std::true_type{});
      else
        ccs(std::false_type{});
    } else if (dbl && o.A32 && o.hash_r0)
      hipLaunchKernelGGL((bi_matvec_kernel<double, MODE, float, true>), dim3(nact), dim3(BI_NT), 0, sm, active, blk,
                         cs, st, *o.A32, Vd, part, nslots, nullptr, nullptr, nullptr, nullptr, early);
    else if (dbl && o.A32)
      hipLaunchKernelGGL((bi_matvec_kernel<double, MODE, float>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs,
                         st, *o.A32, Vd, part, nslots, nullptr, nullptr, nullptr, nullptr, early);
    else
      hipLaunchKernelGGL((bi_matvec_kernel<T, MODE>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, A, V,
                         part, nslots, nullptr, nullptr, nullptr, nullptr, early);
  };
  using M0 = std::integral_constant<int, 0>;
  using M1 = std::integral_constant<int, 1>;
  // the coarse correction of p_ / s_ follows P / S; with bit 8 (stage timing) it runs only as
  // its own stage, bits 6 / 7
  // additive: after P / S from the block sums that kernel wrote; multiplicative: before P / S from
  // the block sums of r, p, q the previous kernels wrote
  auto coarse = [&](int mode) {
    if (!(dbl && o.gs == BiOpts::GSC && o.cinv) && !gsc_mult) return;
    if (o.mult)  // mode 1 forms alpha from the Q partials (no alpha finalisation)
      launch_coarse_apply_mul(sm, active, nact, blk, cs, st, o.cinv, o.coff, o.ce, nslots, mode,
                              mode == 1 ? part0 : nullptr, count, mode == 0 && (mask & 512) ? part0 : nullptr);
    else
      launch_coarse_apply(sm, active, nact, blk, cs, st, o.cinv, o.coff, part + 4 * nslots, o.ce + mode * nslots);
  };
  const bool split = mask & 256, pre = o.mult;
  if ((mask & 1) && !split && pre) coarse(0);
  if (mask & 1) precond(M0{});
  if ((mask & 1) && !split && !pre) coarse(0);
  if (split && (mask & 64)) coarse(0);
  const bool alpha_in_coarse = gsc_mult;
  if (mask & 2) {
    matvec(M0{});
    if (!(mask & 32) && !alpha_in_coarse) launch_finalize(sm, caseact, ncase, cs, st, part, nslots, FIN_BI_ALPHA, count);
  }
  if ((mask & 4) && !split && pre) coarse(1);
  if (mask & 4) precond(M1{});
  if ((mask & 4) && !split && !pre) coarse(1);
  if (split && (mask & 128)) coarse(1);
  if (mask & 8) matvec(M1{});  // gamma is formed by XR from the T partials
  if (mask & 16) {
    auto xr = [&](auto hr_tag) {
      constexpr bool HR = decltype(hr_tag)::value;
      const double* tpart = part + 2 * nslots;
      if (o.cinv && o.mult)  // p_ / s_ already hold P e; block sum of r kept
        hipLaunchKernelGGL((bi_xr_kernel<T, HR>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, x, V, part,
                           nslots, count, nullptr, o.ce + 2 * nslots, tpart, early);
      else if (o.cinv)
        hipLaunchKernelGGL((bi_xr_kernel<T, HR, true>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, x, V,
                           part, nslots, count, (const double*)o.ce, nullptr, tpart, early);
      else
        hipLaunchKernelGGL((bi_xr_kernel<T, HR>), dim3(nact), dim3(BI_NT), 0, sm, active, blk, cs, st, x, V, part,
                           nslots, count, nullptr, nullptr, tpart, early);
    };
    if (o.hash_r0)
      xr(std::true_type{});
    else
      xr(std::false_type{});
    if (!(mask & 32) && !(mask & 1024)) launch_finalize(sm, caseact, ncase, cs, st, part, nslots, FIN_BI_XR, count);
  }
}

#define INST(T)                                                                                                     \
  template void launch_bicgstab_init<T>(hipStream_t, const int*, int, const DevBlock*, const DevCase*, SolveState*,  \
                                        Coef7<T>, const T*, const T*, const T*, BiVecs<T>, double*, size_t, bool,    \
                                        double*);                                                                    \
  template void launch_bicgstab_iter<T>(hipStream_t, const int*, int, const int*, int, const DevBlock*,            \
                                        const DevCase*, SolveState*, Coef7<T>, T*, BiVecs<T>, double*, size_t, int,  \
                                        int, const BiOpts&);                                                         \
  template void launch_bicgstab_reanchor<T>(hipStream_t, const int*, int, const DevBlock*, const SolveState*,         \
                                            BiVecs<T>, double*, size_t);                                             \
  template void launch_bicgstab_restart<T>(hipStream_t, const int*, int, const DevBlock*, const SolveState*, BiVecs<T>); \
  template void launch_bicgstab_shadow<T>(hipStream_t, const int*, int, const DevBlock*, const DevCase*,              \
                                          const SolveState*, BiVecs<T>);
INST(double)
INST(float)
#undef INST

}  // namespace ffvhc
