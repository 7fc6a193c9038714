// Per-case finalisation of the solver reductions.  Every solver kernel writes one partial sum
// per block (two for kernels with two dot products) into part[q * nslots + slot]; this kernel,
// one workgroup per selected case, adds a case's partials in slot order with a fixed tree and
// applies the scalar step of the algorithm (alpha, beta, gamma, IsConverged, ...).  Keeping the
// cross-block step in its own launch leaves the streaming kernels free of device-scope atomics
// and fences, which serialise across the chip's L2 caches.
#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

constexpr int FIN_NT = 256;

__global__ __launch_bounds__(FIN_NT) void finalize_kernel(const int* __restrict__ caseact,
                                                          const DevCase* __restrict__ cs, SolveState* __restrict__ st,
                                                          const double* __restrict__ part, size_t nslots, int op,
                                                          int arg, const SolveState* __restrict__ gate) {
  __shared__ double red[2 * (FIN_NT / 64)];
  const int c = caseact[blockIdx.x];
  if (op == FIN_BI_REANCHOR) {
    if (gate && gate[c].done) return;
  } else if (st[c].done) {
    return;
  }
  const DevCase& C = cs[c];
  const bool two = op == FIN_BI_INIT || op == FIN_BI_GAMMA || op == FIN_BI_XR || op == FIN_CG_INIT ||
                   op == FIN_CG_XR || ((op == FIN_RBGS || op == FIN_FIXAVG) && arg == 1) ||
                   (op == FIN_DEFECT && arg == 0) || op == FIN_BI_REANCHOR;
  double t0 = 0.0, t1 = 0.0;
  if (two) case_sum2<FIN_NT>(part, part + nslots, C.slot_begin, C.nblk, red, t0, t1);
  else t0 = case_sum<FIN_NT>(part, C.slot_begin, C.nblk, red);
  const double t2 = op == FIN_FIXAVG ? case_sum<FIN_NT>(part + 2 * nslots, C.slot_begin, C.nblk, red) : 0.0;
  if (threadIdx.x != 0) return;
  SolveState& S = st[c];
  switch (op) {
    case FIN_BI_INIT:  // FFVILS.h:810-834: t0 = b.b, t1 = r.r0 of the first iteration
      S.bb = t0;
      S.rr0 = 1.0;
      S.alpha = 0.0;
      S.gamma = 1.0;
      if (fabs(t0) < FLT_MIN) {
        S.residual = 0.0;
        S.count = 0;
        S.done = 1;
      } else if (fabs(t1) < FLT_MIN) {
        S.residual = t1;
        S.count = 0;
        S.done = 1;
      } else {
        S.rr1 = t1;
      }
      break;
    case FIN_BI_ALPHA:  // FFVILS.h:868-872: alpha = rr1 / (q_.r0)
      S.alpha = S.rr1 / t0;
      S.restart = 0;
      break;
    case FIN_BI_GAMMA:  // FFVILS.h:905-915: gamma = (t_.s) / (t_.t_)
      S.gamma = t0 / t1;
      break;
    case FIN_BI_XR: {  // FFVILS.h:925-965: t0 = r.r (IsConverged), t1 = next r.r0
      const int n = arg > 0 ? arg : S.iter + 1;  // arg 0: count kept on the device
      const double rho_n = S.rr1;
      fin_bi_xr_store(S, fin_bi_xr_eval(S, t0, t1, n, rho_n), n, rho_n);
      break;
    }
    case FIN_CG_INIT:  // FFVILS.h:679-722: t0 = b.b, t1 = r.z
      S.bb = t0;
      S.rr0 = 1.0;
      if (fabs(t0) < FLT_MIN) {
        S.residual = 0.0;
        S.count = 0;
        S.done = 1;
      } else if (fabs(t1) < FLT_MIN) {
        S.residual = t1;
        S.count = 0;
        S.done = 1;
      } else {
        S.rr1 = t1;
        S.beta = t1 / S.rr0;
      }
      break;
    case FIN_CG_ALPHA:  // FFVILS.h:745-749: alpha = rr1 / (q.p)
      S.alpha = S.rr1 / t0;
      break;
    case FIN_CG_XR:  // FFVILS.h:759-787: t0 = r.r (IsConverged), t1 = next r.z
      S.iter = arg;
      converged(S, t0, arg);
      if (!S.done) {
        S.rr0 = S.rr1;
        if (fabs(t1) < FLT_MIN) {
          S.residual = t1;
          S.count = 0;
          S.done = 1;
        } else {
          S.rr1 = t1;
          S.beta = t1 / S.rr0;
        }
      }
      break;
    case FIN_RBGS:  // launch arg = n computed ||b - A x^{n-1}||^2 (and b.b when n == 1)
      S.iter = arg;
      if (arg == 1) {
        S.bb = t1;
        if (S.countMax < 1) S.done = 1;
      } else {
        converged(S, t0, arg - 1);
        S.parity = (arg - 1) & 1;
      }
      break;
    case FIN_FIXAVG:  // as FIN_RBGS, plus the shift a = -sum(x^n) / Ncells for x^n
      S.iter = arg;
      if (arg == 1) {
        S.bb = t1;
        if (S.countMax < 1) S.done = 1;
      } else {
        converged(S, t0, arg - 1);
        S.parity = (arg - 1) & 1;
      }
      if (!S.done) S.alpha = -t2 / (double)((long long)C.nblk * NC);
      break;
    case FIN_BI_REANCHOR:  // reliable update: t0 = r.r (new reference), t1 = r.r0 (new rho)
      S.bb = t0;
      S.done = 0;
      S.status = 0;
      if (fabs(t1) < FLT_MIN || !(fabs(t1) <= DBL_MAX)) {
        S.restart = 1;  // r0 := r is done by the host; p := r on the next P step
        S.rr1 = t0;
      } else {
        S.rr1 = t1;
        S.beta = t1 / S.rr0 * S.alpha / S.gamma;
      }
      break;
    case FIN_DEFECT:  // outer pass arg: FP64 ||b - A x||^2 (and b.b on the first pass)
      if (arg == 0) S.bb = t1;
      S.iter = arg;
      converged(S, t0, arg);
      break;
    default:
      break;
  }
}

void launch_finalize(hipStream_t sm, const int* caseact, int ncase, const DevCase* cs, SolveState* st,
                     const double* part, size_t nslots, int op, int arg, const SolveState* gate) {
  if (ncase > 0)
    hipLaunchKernelGGL(finalize_kernel, dim3(ncase), dim3(FIN_NT), 0, sm, caseact, cs, st, part, nslots, op, arg,
                       gate);
}

}  // namespace ffvhc
