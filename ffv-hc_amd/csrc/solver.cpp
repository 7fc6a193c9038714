// Linear solvers behind the C ABI (FFVILS.h): fused device-resident RBGS and Jacobi-PCG, and the
// FP32-inner / FP64-defect-correction mixed-precision drivers.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <rocblas/rocblas.h>

#include "kernels.h"
#include "runtime.h"

using namespace ffvhc;

namespace {

// Iterations launched between host checks of the per-case `done` flags; the device skips the
// launches of finished cases, so a batch only bounds how long the host waits.
constexpr int kBatchMin = 8;
constexpr int kBatchMax = 256;

// BiCGSTAB batches without a hint grow 8, 8, 16, 16, 32, ... (doubling every second batch): the
// Krylov solves need tens to hundreds of iterations, and doubling every batch launched up to as many
// iterations after convergence as before it (8 + 16 + 32 for 28 needed) where an extra host check
// costs about one launched iteration
int next_bicg_batch(int batch, int nbatches) { return nbatches % 2 == 0 ? std::min(batch * 2, kBatchMax) : batch; }

// scratch roles
enum { R_R = 1, R_P = 2, R_Q = 3, R_E = 4, R_RHS = 5, R_A = 10, R_BI = 20, R_REC = 30 };

// Orders c->sm after a coarse inversion still running on c->sc (a no-op when none is).
int coarse_join(ffvhc_ctx* c) {
  if (!c->coarse_pending) return 0;
  c->coarse_pending = false;
  HIPCHK(hipStreamWaitEvent(c->sm, c->ev_cd, 0));
  return 0;
}

// Joins the coarse inversion when the entry point that started it returns, on every path.
struct CoarseJoin {
  ffvhc_ctx* c;
  ~CoarseJoin() { (void)coarse_join(c); }
};

template <class T>
Coef7<T> coef(ffvhc_ctx* c, const int* A) {
  Coef7<T> k;
  for (int q = 0; q < 7; q++) k.p[q] = (const T*)c->fields[A[q]].d[0];
  return k;
}

int read_state(ffvhc_ctx* c, SolveState* dst, std::vector<SolveState>& hst, const std::vector<int>& ids,
               bool* all_done) {
  HIPCHK(hipMemcpyAsync(hst.data(), dst, sizeof(SolveState) * c->cases.size(), hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  *all_done = true;
  for (int id : ids)
    if (!hst[id].done) *all_done = false;
  return 0;
}

int check_args(ffvhc_ctx* c, int x, const int* A, int b, int prec) {
  if (!c || !A) return fail(-1, "null argument");
  if (!field_ok(c, x, prec) || !field_ok(c, b, prec)) return fail(-1, "x/b must be fields of one precision");
  for (int q = 0; q < 7; q++)
    if (!field_ok(c, A[q], prec)) return fail(-1, "A bands must be fields of the same precision");
  return 0;
}

// per-case scalar step after a reducing kernel (sums the per-block partials in slot order)
inline void fin(ffvhc_ctx* c, SolveState* st, int op, int arg, const SolveState* gate = nullptr) {
  launch_finalize(c->sm, c->dcaseact, c->ncaseact, c->dcase, st, c->dpart, (size_t)c->cap, op, arg, gate);
}

// RBGS (mode 0), Jacobi (mode 1) or RBGS with fixed average (mode 2) on field x (ping-pong),
// matrix K, rhs bv; state in st.  Result left in buffer 0.
template <class T>
int rbgs_core(ffvhc_ctx* c, int xf, Coef7<T> K, const T* bv, int bcg, double omega, int countMax, double eps,
              SolveState* st, std::vector<SolveState>& hst, const SolveState* outer, int* launched, int mode = 0) {
  const std::vector<int> ids = selected_cases(c);
  int rc0 = ensure_second_buffer(c, xf);
  if (rc0) return rc0;
  const FieldRec X = c->fields[xf];  // copy: no reference into the field table is held
  launch_solve_init(c->sm, st, (int)c->cases.size(), omega, eps, countMax, outer);
  T* xs[2] = {(T*)X.d[0], (T*)X.d[1]};
  T* hs[2] = {(T*)X.h[0], (T*)X.h[1]};
  int n = 1, batch = kBatchMin;
  bool done = false;
  while (!done) {
    int last = std::min(n + batch - 1, countMax + 1);
    for (; n <= last; n++)
    {
      launch_rbgs_iter<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, bv, xs[(n - 1) & 1], xs[n & 1],
                          hs[(n - 1) & 1], hs[n & 1], c->dpart, (size_t)c->cap, n, bcg, mode);
      fin(c, st, mode == 2 ? FIN_FIXAVG : FIN_RBGS, n);
      if (mode == 2)
        launch_shift_push<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, xs[n & 1], hs[(n - 1) & 1],
                             hs[n & 1], bcg);
    }
    HIPCHK(hipGetLastError());
    int rc = read_state(c, st, hst, ids, &done);
    if (rc) return rc;
    if (n > countMax + 1) done = true;
    batch = std::min(batch * 2, kBatchMax);
  }
  if (launched) *launched = n - 1;
  launch_select<T>(c->sm, c->dactive, c->nactive, c->dblk, st, xs[0], xs[1], hs[0], hs[1]);
  HIPCHK(hipGetLastError());
  return 0;
}

// Jacobi-PCG (PreCount = 1) on x, matrix K, rhs bv; scratch r,p,q of the same precision.
template <class T>
int cg_core(ffvhc_ctx* c, int xf, Coef7<T> K, const T* bv, int bcg, double omega, int countMax, double eps,
            SolveState* st, std::vector<SolveState>& hst, const SolveState* outer, int* launched) {
  const std::vector<int> ids = selected_cases(c);
  const int prec = (int)sizeof(T);
  int fr, fp, fq;
  int rc = scratch_field(c, prec, R_R, &fr);
  if (!rc) rc = scratch_field(c, prec, R_P, &fp);
  if (!rc) rc = scratch_field(c, prec, R_Q, &fq);
  if (rc) return rc;
  // copies taken after every allocation
  const FieldRec X = c->fields[xf], Rf = c->fields[fr], Pf = c->fields[fp], Qf = c->fields[fq];
  launch_solve_init(c->sm, st, (int)c->cases.size(), omega, eps, countMax, outer);
  launch_cg_init<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, bv, (const T*)X.d[0],
                    (const T*)X.h[0], (T*)Rf.d[0], c->dpart, (size_t)c->cap);
  fin(c, st, FIN_CG_INIT, 0);
  int n = 1, batch = kBatchMin;
  bool done = false;
  while (!done) {
    int last = std::min(n + batch - 1, countMax);
    for (; n <= last; n++) {
      launch_cg_p<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, (const T*)Rf.d[0], (T*)Pf.d[0],
                     (T*)Pf.h[0], n, BC_NULL);
      launch_cg_q<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, (const T*)Pf.d[0], (const T*)Pf.h[0],
                     (T*)Qf.d[0], c->dpart);
      fin(c, st, FIN_CG_ALPHA, n);
      launch_cg_xr<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, (T*)X.d[0], (T*)Rf.d[0],
                      (const T*)Pf.d[0], (const T*)Qf.d[0], c->dpart, (size_t)c->cap, n);
      fin(c, st, FIN_CG_XR, n);
    }
    HIPCHK(hipGetLastError());
    rc = read_state(c, st, hst, ids, &done);
    if (rc) return rc;
    if (n > countMax) done = true;
    batch = std::min(batch * 2, kBatchMax);
  }
  if (launched) *launched = n - 1;
  // FFVILS.h:789 plsx->ImposeBoundaryCondition after the loop
  launch_ghost_pull<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const T*)X.d[0], (T*)X.h[0], bcg, 1,
                       c->multilevel);
  HIPCHK(hipGetLastError());
  return 0;
}

// BiCGSTAB (PreCount = 1) on x, matrix K, rhs bv; eight scratch vectors of the same precision.
template <class T>
int bicg_vectors(ffvhc_ctx* c, BiVecs<T>* V) {
  const int prec = (int)sizeof(T);
  int f[8];
  for (int q = 0; q < 8; q++) {
    int rc = scratch_field(c, prec, R_BI + q, &f[q]);
    if (rc) return rc;
  }
  T** vp[8] = {&V->r, &V->r0, &V->p, &V->pp, &V->q, &V->s, &V->ss, &V->t};
  for (int q = 0; q < 8; q++) *vp[q] = (T*)c->fields[f[q]].d[0];
  V->hpp = (T*)c->fields[f[3]].h[0];
  V->hss = (T*)c->fields[f[6]].h[0];
  return 0;
}

template <class T>
int bicgstab_core(ffvhc_ctx* c, int xf, Coef7<T> K, const T* bv, int bcg, double omega, int countMax,
                  double eps, SolveState* st, std::vector<SolveState>& hst, const SolveState* outer, int* launched,
                  const BiOpts& opts = BiOpts()) {
  const std::vector<int> ids = selected_cases(c);
  BiVecs<T> V;
  int rc0 = bicg_vectors<T>(c, &V);
  if (rc0) return rc0;
  const FieldRec X = c->fields[xf];  // copy taken after the scratch allocations
  launch_solve_init(c->sm, st, (int)c->cases.size(), omega, eps, countMax, outer);
  launch_bicgstab_init<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, bv, (const T*)X.d[0],
                          (const T*)X.h[0], V, c->dpart, (size_t)c->cap, opts.hash_r0,
                          opts.mult ? opts.ce + 2 * c->cap : nullptr);
  fin(c, st, FIN_BI_INIT, 0);
  // the iterations read the coarse operator of the two-level preconditioners
  int rcj = coarse_join(c);
  if (rcj) return rcj;
  // One iteration's launches (count 0: the iteration number lives in the case state) captured
  // once per solve and replayed: one graph launch per iteration instead of up to ten kernel launches.
  hipGraphExec_t exec = nullptr;
  if (c->graphs) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->sm, hipStreamCaptureModeThreadLocal));
    launch_bicgstab_iter<T>(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase, st, K,
                            (T*)X.d[0], V, c->dpart, (size_t)c->cap, 0, 31, opts);
    const hipError_t ec = hipStreamEndCapture(c->sm, &graph);
    if (ec != hipSuccess || !graph) return fail(-2, std::string("graph capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(-2, std::string("graph instantiate: ") + hipGetErrorString(ei));
  }
  // a round of a repeated mixed solve starts with one batch of the iterations that round needed
  // last time (plus one), then continues from the smallest batch; the device skips the launches
  // of finished cases either way, so the batches change only the launches and host waits
  const int hround = c->bicg_round >= 0 ? std::min(c->bicg_round, 3) : -1;
  int* hint = hround >= 0 ? &c->bicg_hints[c->bicg_hint_key][hround] : nullptr;  // zero when new
  bool hinted = hint && *hint > 0;
  int n = 1, batch = hinted ? std::min(*hint + 1, kBatchMax) : kBatchMin, nbatches = 0;
  bool done = false;
  // multiplicative two-level methods: inside a batch, each iteration's finalisation runs in the
  // next iteration's first coarse application; the batch's last iteration has its own launch, so
  // the state the host reads after the batch is complete
  const bool fuse = c->fuse_fin && !exec && opts.gs == BiOpts::GSC && opts.cinv && opts.mult;
  while (!done) {
    int last = std::min(n + batch - 1, countMax);
    const int first = n;
    for (; n <= last; n++) {
      if (exec) {
        const hipError_t e = hipGraphLaunch(exec, c->sm);
        if (e != hipSuccess) {
          (void)hipGraphExecDestroy(exec);
          return fail(-2, std::string("graph launch: ") + hipGetErrorString(e));
        }
      } else {
        const int mask = 31 | (fuse && n > first ? 512 : 0) | (fuse && n < last ? 1024 : 0);
        launch_bicgstab_iter<T>(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase, st, K,
                                (T*)X.d[0], V, c->dpart, (size_t)c->cap, n, mask, opts);
      }
    }
    HIPCHK(hipGetLastError());
    int rc = read_state(c, st, hst, ids, &done);
    if (rc) {
      if (exec) (void)hipGraphExecDestroy(exec);
      return rc;
    }
    if (n > countMax) done = true;
    batch = hinted ? kBatchMin : next_bicg_batch(batch, ++nbatches);
    if (hinted) nbatches = 0;  // after the hinted batch: 8, 8, 16, 16, ...
    hinted = false;
  }
  if (exec) HIPCHK(hipGraphExecDestroy(exec));
  if (launched) *launched = n - 1;
  if (hint) {
    int m = 0;
    for (int id : ids) m = std::max(m, hst[id].count);
    *hint = m;
  }
  launch_ghost_pull<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const T*)X.d[0], (T*)X.h[0], bcg, 1,
                       c->multilevel);
  HIPCHK(hipGetLastError());
  return 0;
}

int report(ffvhc_ctx* c, const std::vector<SolveState>& hst, int* count, double* residual) {
  const std::vector<int> ids = selected_cases(c);
  int status = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const SolveState& s = hst[ids[i]];
    if (count) count[i] = s.count;
    if (residual) residual[i] = s.residual;
    if (s.status) status = 1;
  }
  if (status) set_error("residual is NaN or Inf");
  return status;
}

int sync_state(ffvhc_ctx* c, SolveState* dst, std::vector<SolveState>& hst) {
  HIPCHK(hipStreamSynchronize(c->sm));
  HIPCHK(hipMemcpy(hst.data(), dst, sizeof(SolveState) * c->cases.size(), hipMemcpyDeviceToHost));
  return 0;
}

template <class T>
int rbgs_fixed_run(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int nsweep, double* ms) {
  int rc = ensure_second_buffer(c, x);
  if (rc) return rc;
  const FieldRec X = c->fields[x];
  launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, -1.0, nsweep);
  Coef7<T> K = coef<T>(c, A);
  const T* bv = (const T*)c->fields[b].d[0];
  T* xs[2] = {(T*)X.d[0], (T*)X.d[1]};
  T* hs[2] = {(T*)X.h[0], (T*)X.h[1]};
  std::vector<hipEvent_t> ev(2 * nsweep + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], c->sm));
  for (int n = 1; n <= nsweep; n++) {
    launch_rbgs_iter<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, bv, xs[(n - 1) & 1], xs[n & 1],
                        hs[(n - 1) & 1], hs[n & 1], c->dpart, (size_t)c->cap, n, bcg);
    HIPCHK(hipEventRecord(ev[2 * n - 1], c->sm));
    fin(c, c->dst, FIN_RBGS, n);
    HIPCHK(hipEventRecord(ev[2 * n], c->sm));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(ev[2 * nsweep]));
  double tk = 0.0, tf = 0.0;
  for (int n = 1; n <= nsweep; n++) {
    float a = 0.f, b2 = 0.f;
    HIPCHK(hipEventElapsedTime(&a, ev[2 * n - 2], ev[2 * n - 1]));
    HIPCHK(hipEventElapsedTime(&b2, ev[2 * n - 1], ev[2 * n]));
    tk += a;
    tf += b2;
  }
  if (ms) {
    ms[0] = tk;
    ms[1] = tf;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (nsweep & 1) {  // the result of sweep nsweep is in buffer (nsweep & 1)
    launch_copy<T>(c->sm, c->dactive, c->nactive, xs[0], hs[0], xs[1], hs[1]);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->sm));
  return 0;
}

// niter PCG iterations with events between the three kernels: ms[0..2] = summed device time
// of the P, Q and XR kernels.
template <class T>
int cg_timing_run(ffvhc_ctx* c, int x, const int* A, int b, double omega, int niter, double* ms) {
  const int prec = (int)sizeof(T);
  int fr, fp, fq;
  int rc = scratch_field(c, prec, R_R, &fr);
  if (!rc) rc = scratch_field(c, prec, R_P, &fp);
  if (!rc) rc = scratch_field(c, prec, R_Q, &fq);
  if (rc) return rc;
  const FieldRec X = c->fields[x], Rf = c->fields[fr], Pf = c->fields[fp], Qf = c->fields[fq];
  Coef7<T> K = coef<T>(c, A);
  launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, -1.0, niter);
  launch_cg_init<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, (const T*)c->fields[b].d[0],
                    (const T*)X.d[0], (const T*)X.h[0], (T*)Rf.d[0], c->dpart, (size_t)c->cap);
  fin(c, c->dst, FIN_CG_INIT, 0);
  std::vector<hipEvent_t> ev(6 * niter);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int n = 1; n <= niter; n++) {
    hipEvent_t* e = &ev[6 * (n - 1)];
    HIPCHK(hipEventRecord(e[0], c->sm));
    launch_cg_p<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, (const T*)Rf.d[0], (T*)Pf.d[0],
                   (T*)Pf.h[0], n, BC_NULL);
    HIPCHK(hipEventRecord(e[1], c->sm));
    launch_cg_q<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, (const T*)Pf.d[0], (const T*)Pf.h[0],
                   (T*)Qf.d[0], c->dpart);
    HIPCHK(hipEventRecord(e[2], c->sm));
    fin(c, c->dst, FIN_CG_ALPHA, n);
    HIPCHK(hipEventRecord(e[3], c->sm));
    launch_cg_xr<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, (T*)X.d[0], (T*)Rf.d[0],
                    (const T*)Pf.d[0], (const T*)Qf.d[0], c->dpart, (size_t)c->cap, n);
    HIPCHK(hipEventRecord(e[4], c->sm));
    fin(c, c->dst, FIN_CG_XR, n);
    HIPCHK(hipEventRecord(e[5], c->sm));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->sm));
  for (int k = 0; k < 4; k++) ms[k] = 0.0;
  static const int slot_of[5] = {0, 1, 3, 2, 3};  // p, q, fin, xr, fin
  for (int n = 0; n < niter; n++)
    for (int k = 0; k < 5; k++) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev[6 * n + k], ev[6 * n + k + 1]));
      ms[slot_of[k]] += t;
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return 0;
}

// niter BiCGSTAB iterations with events between its five kernels; ms[k] = summed device time.
template <class T>
int bicg_timing_run(ffvhc_ctx* c, int x, const int* A, int b, double omega, int niter, double* ms,
                    const BiOpts& o = BiOpts()) {
  BiVecs<T> V;
  int rc = bicg_vectors<T>(c, &V);
  if (rc) return rc;
  const FieldRec X = c->fields[x];
  Coef7<T> K = coef<T>(c, A);
  launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, -1.0, niter);
  launch_bicgstab_init<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K,
                          (const T*)c->fields[b].d[0], (const T*)X.d[0], (const T*)X.h[0], V, c->dpart,
                          (size_t)c->cap, o.hash_r0, o.mult ? o.ce + 2 * c->cap : nullptr);
  fin(c, c->dst, FIN_BI_INIT, 0);
  // stages in launch order: P, coarse(p_), Q, fin(alpha), S, coarse(s_), T, (gamma: formed inside
  // XR, an empty stage), XR, fin(convergence, beta); the coarse applies (two-level methods only)
  // time as their own stage (the multiplicative coarse correction runs before P / S instead, and
  // its s_ application forms alpha, so those methods have no alpha finalisation stage)
  constexpr int NS = 10;
  static const int bit_add[NS] = {1, 64, 2, 0, 4, 128, 8, 0, 16, 0};
  static const int slot_add[NS] = {0, 6, 1, 5, 2, 6, 3, 5, 4, 5};
  static const int bit_mul[NS] = {64, 1, 2, 0, 128, 4, 8, 0, 16, 0};
  static const int slot_mul[NS] = {6, 0, 1, 5, 6, 2, 3, 5, 4, 5};
  static const int fin_add[NS] = {-1, -1, -1, FIN_BI_ALPHA, -1, -1, -1, -1, -1, FIN_BI_XR};
  static const int fin_mul[NS] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, FIN_BI_XR};  // alpha: coarse(s_)
  const bool mul_alpha = o.mult && o.gs == BiOpts::GSC && o.cinv;
  const int* fin_of = mul_alpha ? fin_mul : fin_add;
  const int* bit_of = o.mult ? bit_mul : bit_add;
  const int* slot_of = o.mult ? slot_mul : slot_add;
  std::vector<hipEvent_t> ev((NS + 1) * niter);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int n = 1; n <= niter; n++) {
    hipEvent_t* e = &ev[(NS + 1) * (n - 1)];
    HIPCHK(hipEventRecord(e[0], c->sm));
    for (int k = 0; k < NS; k++) {
      if (bit_of[k])
        launch_bicgstab_iter<T>(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase, c->dst,
                                K, (T*)X.d[0], V, c->dpart, (size_t)c->cap, n, bit_of[k] | 32 | 256, o);
      else if (fin_of[k] >= 0)
        fin(c, c->dst, fin_of[k], n);
      HIPCHK(hipEventRecord(e[k + 1], c->sm));
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->sm));
  for (int k = 0; k < 7; k++) ms[k] = 0.0;
  for (int n = 0; n < niter; n++)
    for (int k = 0; k < NS; k++) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev[(NS + 1) * n + k], ev[(NS + 1) * n + k + 1]));
      ms[slot_of[k]] += t;
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return 0;
}

// Panel width of the GEMM Gauss-Jordan: each panel's trailing update is a rank-GJW FP64 GEMM
constexpr int GJW = 256;

// doubles of scratch the Gauss-Jordan of an n x n matrix needs: the 32-wide form GJB * n; the wide
// form D^-1 (GJW^2), the 32-wide scratch of D's own inversion, the old panel columns C (n x GJW)
// and the new panel rows T (GJW x n)
size_t gj_scratch(ffvhc_ctx* c, size_t n) {
  if ((int)n < c->coarse_gemm_min) return 32 * n;
  return (size_t)GJW * GJW + 32 * (size_t)GJW + 2 * (size_t)GJW * n;
}

#define BLASCHK(expr)                                                                     \
  do {                                                                                    \
    rocblas_status s_ = (expr);                                                           \
    if (s_ != rocblas_status_success)                                                     \
      return fail(-2, std::string(#expr) + ": " + rocblas_status_to_string(s_));          \
  } while (0)

// In-place inverse of the row-major n x n matrix M (no pivoting: A_c is an M-matrix-like operator
// with a dominant diagonal) by Gauss-Jordan over panels K of GJW columns.  Per panel, with D = M_KK
// and I, J the indices outside K:
//   D^-1 (D copied out and inverted by the 32-wide Gauss-Jordan), C = M_.K with the rows of K zeroed,
//   T = D^-1 M_K.  with the columns of K zeroed   (GEMM, GJW x GJW x n)
//   M_IJ -= C_I T_J                               (GEMM, n x n x GJW; the zero rows / columns leave
//                                                  the rows and columns of K alone)
//   M_.K = -C D^-1                                (GEMM, n x GJW x GJW; rows of K come out 0)
//   M_K. = T, M_KK = D^-1
// rocBLAS is column-major: a row-major matrix is its transpose there, so every product is issued
// with its operands swapped.
int gauss_jordan_inverse_gemm(ffvhc_ctx* c, hipStream_t sm, double* M, int n, double* work) {
  if (n <= 0) return 0;
  if (!c->blas) {
    rocblas_handle h = nullptr;
    BLASCHK(rocblas_create_handle(&h));
    c->blas = h;
  }
  rocblas_handle h = (rocblas_handle)c->blas;
  BLASCHK(rocblas_set_stream(h, sm));
  BLASCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
  double* D = work;
  double* Ts = D + (size_t)GJW * GJW;
  double* C = Ts + 32 * (size_t)GJW;
  double* T = C + (size_t)GJW * n;
  const double one = 1.0, zero = 0.0, mone = -1.0;
  for (int k0 = 0; k0 < n; k0 += GJW) {
    const int kb = std::min(GJW, n - k0);
    launch_gjw_get_diag(sm, M, n, k0, kb, D);
    gauss_jordan_inverse(sm, D, kb, Ts);
    launch_gjw_get_cols(sm, M, n, k0, kb, C);
    // T (row-major kb x n) = D^-1 M_K.  <=>  T^T (n x kb, ld n) = M_K.^T (n x kb, ld n) D^-T
    BLASCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, kb, kb, &one, M + (size_t)k0 * n, n,
                          D, kb, &zero, T, n));
    launch_gjw_zero_tcols(sm, T, n, k0, kb);
    // M -= C T  <=>  M^T (n x n, ld n) -= T^T (n x kb, ld n) C^T (kb x n, ld kb)
    BLASCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, kb, &mone, T, n, C, kb, &one, M,
                          n));
    // M_.K = -C D^-1  <=>  M_.K^T (kb x n, ld n) = -D^-T (kb x kb, ld kb) C^T (kb x n, ld kb)
    BLASCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, kb, n, kb, &mone, D, kb, C, kb, &zero,
                          M + k0, n));
    launch_gjw_put_rows(sm, M, n, k0, kb, T, D);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// Coarse correction of the two-level methods: A_c = P^T A P per selected case (rows from the
// device, assembled densely on the device), inverted on the device (Gauss-Jordan, FP64), kept as
// FP32.  A case whose A_c has zero row sums throughout (no Dirichlet face: A_c singular) has its
// last coarse unknown decoupled, which keeps the coarse correction a fixed linear map.
// First part: the coarse rows of every active block and their readback into c->hcrows, complete
// after the next synchronisation of c->sm.
int coarse_rows_launch(ffvhc_ctx* c, const Coef7<double>& K) {
  c->setup_serial++;
  int fr;
  int rc = scratch_field(c, 8, R_REC, &fr);
  if (rc) return rc;
  double* drows = (double*)c->fields[fr].d[0];  // 25 doubles per slot (cap * NC doubles available)
  launch_coarse_rows(c->sm, c->dactive, c->nactive, c->dblk, K, drows);
  HIPCHK(hipGetLastError());
  c->hcrows.resize((size_t)c->cap * 25);
  HIPCHK(hipMemcpyAsync(c->hcrows.data(), drows, c->hcrows.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c->sm));
  return 0;
}

// Second part, after that synchronisation: assembly and decoupling are queued on c->sm, the
// inversion and the FP32 copy on c->sc after them, and the call returns without waiting (the
// solver's first kernels, which do not read the coarse operator, run beside the inversion);
// coarse_join makes c->sm wait for it before anything reads it.  With ac_out / inv_out (copies
// of A_c / its inverse for the caller) everything runs on c->sm and the call waits.
int coarse_finish(ffvhc_ctx* c, BiOpts* o, double* ac_out, double* inv_out) {
  int fr;
  int rc = scratch_field(c, 8, R_REC, &fr);
  if (rc) return rc;
  rc = coarse_join(c);  // a previous inversion still writing dcm / dci
  if (rc) return rc;
  const bool side = !ac_out && !inv_out;
  if (side && !c->sc) {
    HIPCHK(hipStreamCreateWithFlags(&c->sc, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_cs, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_cd, hipEventDisableTiming));
  }
  const double* drows = (const double*)c->fields[fr].d[0];
  const std::vector<double>& rows = c->hcrows;
  const std::vector<int> ids = selected_cases(c);
  // FP64 operators / inverses at off[case] (n x n); their FP32 copies at offp[case] with the row
  // pitch padded to a multiple of 4 (float4 rows)
  c->hcoff.assign(2 * (size_t)MAX_CASES, 0);
  long long* off = c->hcoff.data();
  long long* offp = off + MAX_CASES;
  size_t total = 0, totalp = 0, maxn = 0;
  for (int id : ids) {
    off[id] = (long long)total;
    offp[id] = (long long)totalp;
    const size_t n = (size_t)c->cases[id].nblk;
    total += n * n;
    totalp += n * ((n + 3) & ~(size_t)3);
    maxn = std::max(maxn, n);
  }
  if (total > c->dcm_n || totalp > c->dci_n) {
    (void)hipFree(c->dcm);
    (void)hipFree(c->dci);
    c->dcm = nullptr;
    c->dci = nullptr;
    HIPCHK(hipMalloc(&c->dcm, total * sizeof(double)));
    HIPCHK(hipMalloc(&c->dci, totalp * sizeof(float)));
    c->dcm_n = total;
    c->dci_n = totalp;
  }
  if (gj_scratch(c, maxn) > c->dcsnap_n) {  // panel scratch of the blocked Gauss-Jordan
    (void)hipFree(c->dcsnap);
    c->dcsnap = nullptr;
    HIPCHK(hipMalloc(&c->dcsnap, gj_scratch(c, maxn) * sizeof(double)));
    c->dcsnap_n = gj_scratch(c, maxn);
  }
  if (!c->dcoff) HIPCHK(hipMalloc(&c->dcoff, 2 * MAX_CASES * sizeof(long long)));
  if (!c->dce) HIPCHK(hipMalloc(&c->dce, 5 * (size_t)c->cap * sizeof(double)));
  HIPCHK(hipMemcpyAsync(c->dcoff, off, 2 * (size_t)MAX_CASES * sizeof(long long), hipMemcpyHostToDevice, c->sm));
  launch_coarse_assemble(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dcoff, drows, c->dcm, total);
  for (int id : ids) {
    // singular A_c (zero row sums throughout: no Dirichlet face) -> decouple the last unknown
    const CaseRec& cr = c->cases[id];
    double dmax = 0.0, smax = 0.0;
    for (int a = 0; a < cr.nblk; a++) {
      const int slot = cr.slot_begin + a;
      const DevBlock& B = c->hblk[slot];
      double rs = rows[(size_t)slot * 25];
      for (int f = 0; f < 6; f++) {
        if (B.ftype[f] == FACE_OUTER) continue;
        const int ns = B.ftype[f] == FACE_FINER ? 4 : 1;
        for (int s = 0; s < ns; s++) {
          rs += rows[(size_t)slot * 25 + 1 + 4 * f + s];
          if (B.nbr[f][s] == slot) dmax = std::max(dmax, std::fabs(rows[(size_t)slot * 25 + 1 + 4 * f + s]));
        }
      }
      dmax = std::max(dmax, std::fabs(rows[(size_t)slot * 25]));
      smax = std::max(smax, std::fabs(rs));
    }
    if (smax <= 1e-12 * dmax) launch_coarse_decouple(c->sm, c->dcm + off[id], cr.nblk);
  }
  if (ac_out && !ids.empty())
    HIPCHK(hipMemcpyAsync(ac_out, c->dcm + off[ids[0]],
                          (size_t)c->cases[ids[0]].nblk * c->cases[ids[0]].nblk * sizeof(double),
                          hipMemcpyDeviceToHost, c->sm));
  hipStream_t gs = c->sm;
  if (side) {
    HIPCHK(hipEventRecord(c->ev_cs, c->sm));
    HIPCHK(hipStreamWaitEvent(c->sc, c->ev_cs, 0));
    gs = c->sc;
  }
  for (int id : ids) {
    const int n = c->cases[id].nblk;
    if (n >= c->coarse_gemm_min) {
      rc = gauss_jordan_inverse_gemm(c, gs, c->dcm + off[id], n, c->dcsnap);
      if (rc) return rc;
    } else {
      gauss_jordan_inverse(gs, c->dcm + off[id], n, c->dcsnap);
    }
  }
  for (int id : ids) {
    const int n = c->cases[id].nblk;
    launch_to_f32(gs, c->dcm + off[id], c->dci + offp[id], n, (n + 3) & ~3);
  }
  HIPCHK(hipGetLastError());
  if (side) {
    HIPCHK(hipEventRecord(c->ev_cd, c->sc));
    c->coarse_pending = true;
  }
  if (inv_out && !ids.empty())
    HIPCHK(hipMemcpyAsync(inv_out, c->dcm + off[ids[0]],
                          (size_t)c->cases[ids[0]].nblk * c->cases[ids[0]].nblk * sizeof(double),
                          hipMemcpyDeviceToHost, c->sm));
  if (ac_out || inv_out) HIPCHK(hipStreamSynchronize(c->sm));  // the caller reads them on return
  o->cinv = c->dci;
  o->coff = c->dcoff + MAX_CASES;
  o->ce = c->dce;
  return 0;
}

int coarse_setup(ffvhc_ctx* c, const Coef7<double>& K, BiOpts* o, double* ac_out = nullptr,
                 double* inv_out = nullptr) {
  int rc = coarse_rows_launch(c, K);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->sm));
  return coarse_finish(c, o, ac_out, inv_out);
}

// Stencil classes of the active blocks for the compact-record methods: the 27 class rows and the
// class-uniform flag of every block (device) and the counts of uniform and other blocks, which
// pick the kernel variant.  FFVHC_OPTION_CLASSES 0 leaves o without them (every block then reads
// the per-cell arrays; the results are the same bit for bit).  class_launch queues the class
// kernel and the readback of the flags into c->hcuni; class_finish, after the next
// synchronisation of c->sm, counts them.
bool class_on(ffvhc_ctx* c) { return c->classes && c->nactive; }

int class_launch(ffvhc_ctx* c, const Coef7<double>& K) {
  c->setup_serial++;
  const size_t cap = (size_t)c->cap;
  if (!c->dcuni) {
    HIPCHK(hipMalloc(&c->dcuni, cap * sizeof(int)));
    HIPCHK(hipMalloc(&c->dctab, cap * 27 * 8 * sizeof(float)));
    HIPCHK(hipMalloc(&c->dctrec, cap * 27 * sizeof(uint4)));
  }
  launch_stencil_classes(c->sm, c->dactive, c->nactive, K, c->dcuni, c->dctab, (uint4*)c->dctrec);
  HIPCHK(hipGetLastError());
  c->hcuni.resize(cap);
  HIPCHK(hipMemcpyAsync(c->hcuni.data(), c->dcuni, cap * sizeof(int), hipMemcpyDeviceToHost, c->sm));
  return 0;
}

void class_finish(ffvhc_ctx* c, BiOpts* o) {
  int nreg = 0;
  for (int s : c->hactive) nreg += c->hcuni[s] != 0;
  o->tab = c->dctab;
  o->trec = (const uint4*)c->dctrec;
  o->uni = c->dcuni;
  o->nreg = nreg;
  o->nirr = c->nactive - nreg;
}

int class_setup(ffvhc_ctx* c, const Coef7<double>& K, BiOpts* o) {
  if (!class_on(c)) return 0;
  int rc = class_launch(c, K);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->sm));
  class_finish(c, o);
  return 0;
}

// Storage options of the FP64-recurrence BiCGSTAB methods 5-9: an FP32 copy of the FP64 matrix A
// (scratch fields, converted here) in *M32 and/or the FP64 matrix in *M64, referenced by *o.
int fp64_bicg_opts(ffvhc_ctx* c, int method, const int* A, Coef7<float>* M32, Coef7<double>* M64, BiOpts* o) {
  c->setup_serial++;
  *o = BiOpts();
  *M64 = coef<double>(c, A);
  if (method == 6) {
    o->gs = BiOpts::GS64;
    o->M64 = M64;
    return 0;
  }
  int fa[7];
  for (int q = 0; q < 7; q++) {
    int rc = scratch_field(c, 4, R_A + q, &fa[q]);
    if (rc) return rc;
  }
  // the compact-record methods (11-28) read no per-cell FP32 band or record when every active
  // block is class-uniform: those arrays are then not built
  bool percell = true;
  if (method >= 11) {
    // the class flags and (methods 17-28) the coarse rows come back to the host in one
    // synchronisation
    const bool cl = class_on(c);
    int rc = cl ? class_launch(c, *M64) : 0;
    if (!rc && method >= 17) rc = coarse_rows_launch(c, *M64);
    if (rc) return rc;
    if (cl || method >= 17) HIPCHK(hipStreamSynchronize(c->sm));
    if (cl) class_finish(c, o);
    percell = !o->trec || o->nirr > 0;
  }
  if (percell)
    for (int q = 0; q < 7; q++)
      launch_convert(c->sm, c->dactive, c->nactive, (float*)c->fields[fa[q]].d[0],
                     (const double*)c->fields[A[q]].d[0]);
  *M32 = coef<float>(c, fa);
  if (method == 5) {
    o->gs = BiOpts::GS32;
    o->M32 = M32;
  } else if (method == 7) {
    o->A32 = M32;
  } else if (method <= 10) {  // 8-10: FP32 matrix storage everywhere, (symmetric) GS in FP64 arithmetic
    o->A32 = M32;
    o->gs = method == 9 ? BiOpts::SGS32_ARITH64 : BiOpts::GS32_ARITH64;
    o->M32 = M32;
    o->hash_r0 = method == 10;
  } else {  // 11-16: FP32 matrix for the matvecs, 1-6 SOR sweeps on the compact records;
            // 17-22: the same plus the additive coarse correction, 23-28 the multiplicative one
    int fr;
    const int rc = scratch_field(c, 16, R_REC, &fr);
    if (rc) return rc;
    o->A32 = M32;
    o->gs = BiOpts::GSC;
    o->hash_r0 = c->shadow_hash;
    o->rec = (const uint4*)c->fields[fr].d[0];
    o->sweeps = (method - 11) % 6 + 1;
    if (percell) launch_gs_pack(c->sm, c->dactive, c->nactive, *M64, (uint4*)c->fields[fr].d[0]);
    o->mult = method >= 23;
    if (method >= 17) return coarse_finish(c, o, nullptr, nullptr);
  }
  return 0;
}

// The kept setup of the time step's pressure solve: valid while no setup buffer was written since
// (setup_serial), the matrix inputs are unchanged (geom_serial) and method, matrix fields, active
// blocks, options and the reference-pressure point match.
bool psetup_match(ffvhc_ctx* c, int method, const int* A) {
  const auto& s = c->psetup;
  if (!c->reuse_setup || !s.valid || s.method != method || s.setup_serial != c->setup_serial ||
      s.geom_serial != c->geom_serial || s.classes != c->classes || s.shadow != c->shadow_hash ||
      s.active != c->hactive || s.ref_active != c->step_ref_active)
    return false;
  if (s.ref_active)
    for (int a = 0; a < 3; a++)
      if (s.ref[a] != c->step_ref[a]) return false;
  for (int q = 0; q < 7; q++)
    if (s.A[q] != A[q]) return false;
  return true;
}

void psetup_store(ffvhc_ctx* c, int method, const int* A, const BiOpts& o) {
  auto& s = c->psetup;
  s.valid = c->reuse_setup;
  s.method = method;
  for (int q = 0; q < 7; q++) s.A[q] = A[q];
  s.setup_serial = c->setup_serial;
  s.geom_serial = c->geom_serial;
  s.classes = c->classes;
  s.shadow = c->shadow_hash;
  s.ref_active = c->step_ref_active;
  for (int a = 0; a < 3; a++) s.ref[a] = c->step_ref[a];
  s.active = c->hactive;
  s.o = o;
  s.o.A32 = nullptr;  // points into the caller's frame; set again on reuse
}

int prologue(ffvhc_ctx* c, int x, const int* A, int b, int bcg, int* prec) {
  if (!c) return fail(-1, "null context");
  *prec = field_ok(c, x, 0) ? c->fields[x].prec : 8;
  int rc = check_args(c, x, A, b, *prec);
  if (rc) return rc;
  if (bcg < 0 || bcg >= NUM_BC) return fail(-1, "bad bc group");
  return refresh_active(c);
}

}  // namespace

extern "C" {

int ffvhc_field_fill(ffvhc_ctx* c, int f, double v) {
  if (!c || !field_ok(c, f, 0)) return fail(-1, "bad field");
  int rc = refresh_active(c);
  if (rc) return rc;
  FieldRec& F = c->fields[f];
  if (F.prec == 8) launch_fill<double>(c->sm, c->dactive, c->nactive, (double*)F.d[0], (double*)F.h[0], v);
  else launch_fill<float>(c->sm, c->dactive, c->nactive, (float*)F.d[0], (float*)F.h[0], (float)v);
  HIPCHK(hipGetLastError());
  return 0;
}

int ffvhc_field_copy(ffvhc_ctx* c, int dst, int src) {
  if (!c || !field_ok(c, dst, 0) || !field_ok(c, src, c->fields[dst].prec)) return fail(-1, "bad fields");
  int rc = refresh_active(c);
  if (rc) return rc;
  FieldRec &D = c->fields[dst], &S = c->fields[src];
  if (D.prec == 8)
    launch_copy<double>(c->sm, c->dactive, c->nactive, (double*)D.d[0], (double*)D.h[0], (const double*)S.d[0],
                        (const double*)S.h[0]);
  else
    launch_copy<float>(c->sm, c->dactive, c->nactive, (float*)D.d[0], (float*)D.h[0], (const float*)S.d[0],
                       (const float*)S.h[0]);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"

namespace {

int stationary(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax, double eps,
               int* count, double* residual, int mode) {
  int prec = 8;
  int rc = prologue(c, x, A, b, bcg, &prec);
  if (rc) return rc;
  if (countMax < 1) {  // for(count=1; count<=countMax; ...) never runs
    const std::vector<int> ids = selected_cases(c);
    for (size_t i = 0; i < ids.size(); i++)
      if (count) count[i] = 1;
    return 0;
  }
  rc = ensure_second_buffer(c, x);
  if (rc) return rc;
  if (prec == 8)
    rc = rbgs_core<double>(c, x, coef<double>(c, A), (const double*)c->fields[b].d[0], bcg, omega, countMax,
                           eps, c->dst, c->hst, nullptr, nullptr, mode);
  else
    rc = rbgs_core<float>(c, x, coef<float>(c, A), (const float*)c->fields[b].d[0], bcg, omega, countMax,
                          eps, c->dst, c->hst, nullptr, nullptr, mode);
  if (rc) return rc;
  rc = sync_state(c, c->dst, c->hst);
  if (rc) return rc;
  return report(c, c->hst, count, residual);
}

}  // namespace

namespace {

// Mixed-precision BiCGSTAB with reliable updates: the Krylov recurrence runs in FP32 on the
// correction e (x = x64 + e); whenever the recursive residual has dropped by delta since the
// last anchor, e is folded into the FP64 solution, the true residual b - A x64 is computed in
// FP64 (the convergence test), and it replaces the FP32 residual without restarting the
// recurrence (rho recomputed; a restart with r0 := r only after a breakdown).
int bicgstab_reliable(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax, double eps,
                      double delta, int* count, double* residual, int* outer, bool random_shadow) {
  const std::vector<int> ids = selected_cases(c);
  int fa[7], fe;
  for (int q = 0; q < 7; q++) {
    int rc = scratch_field(c, 4, R_A + q, &fa[q]);
    if (rc) return rc;
  }
  int rc = scratch_field(c, 4, R_E, &fe);
  if (rc) return rc;
  BiVecs<float> V;
  rc = bicg_vectors<float>(c, &V);
  if (rc) return rc;
  c->setup_serial++;
  for (int q = 0; q < 7; q++)
    launch_convert(c->sm, c->dactive, c->nactive, (float*)c->fields[fa[q]].d[0],
                   (const double*)c->fields[A[q]].d[0]);
  Coef7<double> K64 = coef<double>(c, A);
  Coef7<float> K32 = coef<float>(c, fa);
  const FieldRec X = c->fields[x], E = c->fields[fe];
  const double* b64 = (const double*)c->fields[b].d[0];
  const int max_outer = 100000;
  std::vector<int> outers(c->cases.size(), -1);
  launch_solve_init(c->sm, c->dst2, (int)c->cases.size(), omega, eps, max_outer);
  int o = 0;
  auto defect = [&]() {
    launch_defect(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst2, K64, b64, (const double*)X.d[0],
                  (const double*)X.h[0], V.r, c->dpart, (size_t)c->cap, o);
    fin(c, c->dst2, FIN_DEFECT, o);
  };
  defect();
  bool done = false;
  rc = read_state(c, c->dst2, c->hst2, ids, &done);
  if (rc) return rc;
  for (int id : ids)
    if (c->hst2[id].done) outers[id] = 0;
  int n = 1;
  if (!done) {
    launch_fill<float>(c->sm, c->dactive, c->nactive, (float*)E.d[0], (float*)E.h[0], 0.0f);
    launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, delta, countMax, c->dst2);
    launch_bicgstab_init<float>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K32, V.r,
                                (const float*)E.d[0], (const float*)E.h[0], V, c->dpart, (size_t)c->cap);
    fin(c, c->dst, FIN_BI_INIT, 0);
    if (random_shadow) {  // seeded shadow residual, then rho = r.r0
      launch_bicgstab_shadow<float>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst2, V);
      launch_bicgstab_reanchor<float>(c->sm, c->dactive, c->nactive, c->dblk, c->dst2, V, c->dpart, (size_t)c->cap);
      fin(c, c->dst, FIN_BI_REANCHOR, 0, c->dst2);
    }
    while (true) {
      // one segment: FP32 iterations until the recursive residual dropped by delta
      int batch = kBatchMin;
      bool seg = false;
      while (!seg && n <= countMax) {
        const int last = std::min(n + batch - 1, countMax);
        for (; n <= last; n++)
          launch_bicgstab_iter<float>(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase,
                                      c->dst, K32, (float*)E.d[0], V, c->dpart, (size_t)c->cap, n);
        HIPCHK(hipGetLastError());
        rc = read_state(c, c->dst, c->hst, ids, &seg);
        if (rc) return rc;
        batch = std::min(batch * 2, 64);
      }
      // reliable update: x64 += e, e = 0, FP64 residual
      launch_add_correction(c->sm, c->dactive, c->nactive, c->dblk, c->dst2, c->dst, (double*)X.d[0],
                            (const float*)E.d[0]);
      launch_fill<float>(c->sm, c->dactive, c->nactive, (float*)E.d[0], (float*)E.h[0], 0.0f);
      launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const double*)X.d[0],
                                (double*)X.h[0], bcg, 1, c->multilevel);
      o++;
      defect();
      rc = read_state(c, c->dst2, c->hst2, ids, &done);
      if (rc) return rc;
      for (int id : ids)
        if (c->hst2[id].done && outers[id] < 0) outers[id] = o;
      if (done || n > countMax) break;
      launch_bicgstab_reanchor<float>(c->sm, c->dactive, c->nactive, c->dblk, c->dst2, V, c->dpart, (size_t)c->cap);
      fin(c, c->dst, FIN_BI_REANCHOR, n, c->dst2);
      launch_bicgstab_restart<float>(c->sm, c->dactive, c->nactive, c->dblk, c->dst, V);
      HIPCHK(hipGetLastError());
    }
  }
  rc = sync_state(c, c->dst2, c->hst2);
  if (rc) return rc;
  int status = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const SolveState& s = c->hst2[ids[i]];
    if (count) count[i] = n - 1;
    if (residual) residual[i] = s.residual;
    if (outer) outer[i] = outers[ids[i]] < 0 ? o : outers[ids[i]];
    if (s.status) status = 1;
  }
  if (status) set_error("residual is NaN or Inf");
  return status;
}

}  // namespace

namespace {

// Method 29: FP64 defect correction (iterative refinement) around an FP32 inner solve, the inner solve
// being method 26's BiCGSTAB (FP32 matrix, coarse correction then 4 block SOR sweeps) on FP32
// vectors: each round forms the FP64 residual r = b - A x of the FP64 matrix (and its FP32 copy),
// solves A e = r in FP32 to the relative tolerance inner_eps (at most countMax iterations), and adds
// e to x in FP64; the answer is accepted on the FP64 residual, as for every other method.  The
// inner solve moves half the bytes per iteration of method 26; FP32 limits each round to about
// 1e-6, so eps = 1e-10 takes two or three rounds.
int two_level_defect_correction(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax,
                                double eps, double inner_eps, int* count, double* residual, int* outer) {
  if (!(inner_eps > 0.0 && inner_eps < 1.0)) return fail(-1, "method 29: inner tolerance must be in (0, 1)");
  Coef7<float> M32;
  Coef7<double> K64;
  BiOpts o;
  int rc;
  int fa[7];
  for (int q = 0; q < 7; q++) {
    rc = scratch_field(c, 4, R_A + q, &fa[q]);
    if (rc) return rc;
  }
  if (c->step_solve && psetup_match(c, 29, A)) {  // the time step's matrix this setup was made from
    K64 = coef<double>(c, A);
    M32 = coef<float>(c, fa);
    o = c->psetup.o;
    o.A32 = &M32;
  } else {
    rc = fp64_bicg_opts(c, 26, A, &M32, &K64, &o);
    if (rc) return rc;
    // the inner solve's first residual reads the FP32 bands of every block (class-uniform ones
    // included, which the sweeps and matvecs take from their class rows)
    for (int q = 0; q < 7; q++)
      launch_convert(c->sm, c->dactive, c->nactive, (float*)c->fields[fa[q]].d[0],
                     (const double*)c->fields[A[q]].d[0]);
    if (c->step_solve) psetup_store(c, 29, A, o);
  }
  int fr, fe;
  rc = scratch_field(c, 4, R_RHS, &fr);
  if (!rc) rc = scratch_field(c, 4, R_E, &fe);
  if (rc) return rc;
  const std::vector<int> ids = selected_cases(c);
  const FieldRec X = c->fields[x], E = c->fields[fe], Rf = c->fields[fr];
  std::vector<int> inner_total(c->cases.size(), 0), outers(c->cases.size(), -1);
  const int max_outer = 64;
  launch_solve_init(c->sm, c->dst2, (int)c->cases.size(), omega, eps, max_outer);
  for (int k = 0;; k++) {
    launch_defect(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst2, K64, (const double*)c->fields[b].d[0],
                  (const double*)X.d[0], (const double*)X.h[0], (float*)Rf.d[0], c->dpart, (size_t)c->cap, k);
    fin(c, c->dst2, FIN_DEFECT, k);
    HIPCHK(hipGetLastError());
    bool done = false;
    rc = read_state(c, c->dst2, c->hst2, ids, &done);
    if (rc) return rc;
    for (int id : ids)
      if (c->hst2[id].done && outers[id] < 0) outers[id] = k;
    if (done || k >= max_outer) break;
    launch_fill<float>(c->sm, c->dactive, c->nactive, (float*)E.d[0], (float*)E.h[0], 0.0f);
    c->bicg_round = k;
    c->bicg_hint_key = 29;
    rc = bicgstab_core<float>(c, fe, M32, (const float*)Rf.d[0], BC_HOM, omega, countMax, inner_eps, c->dst, c->hst,
                              c->dst2, nullptr, o);
    c->bicg_round = -1;
    if (rc) return rc;
    rc = sync_state(c, c->dst, c->hst);
    if (rc) return rc;
    for (int id : ids)
      if (!c->hst2[id].done) inner_total[id] += std::max(c->hst[id].count, 0);
    launch_add_correction(c->sm, c->dactive, c->nactive, c->dblk, c->dst2, c->dst, (double*)X.d[0],
                          (const float*)E.d[0]);
    launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const double*)X.d[0],
                              (double*)X.h[0], bcg, 1, c->multilevel);
    HIPCHK(hipGetLastError());
  }
  rc = sync_state(c, c->dst2, c->hst2);
  if (rc) return rc;
  int status = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const SolveState& s = c->hst2[ids[i]];
    if (count) count[i] = inner_total[ids[i]];
    if (residual) residual[i] = s.residual;
    if (outer) outer[i] = outers[ids[i]] < 0 ? max_outer : outers[ids[i]];
    if (s.status) status = 1;
  }
  if (status) set_error("residual is NaN or Inf");
  return status;
}

}  // namespace

extern "C" {

int ffvhc_rbgs(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax, double eps, int* count,
               double* residual) {
  return stationary(c, x, A, b, bcg, omega, countMax, eps, count, residual, 0);
}

int ffvhc_jacobi(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax, double eps,
                 int* count, double* residual) {
  return stationary(c, x, A, b, bcg, omega, countMax, eps, count, residual, 1);
}

int ffvhc_rbgs_fixed_average(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax,
                             double eps, int* count, double* residual) {
  return stationary(c, x, A, b, bcg, omega, countMax, eps, count, residual, 2);
}

int ffvhc_rbgs_fixed(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int nsweep, double* ms) {
  int prec = 8;
  int rc = prologue(c, x, A, b, bcg, &prec);
  if (rc) return rc;
  if (nsweep < 1) return fail(-1, "nsweep must be >= 1");
  return prec == 8 ? rbgs_fixed_run<double>(c, x, A, b, bcg, omega, nsweep, ms)
                   : rbgs_fixed_run<float>(c, x, A, b, bcg, omega, nsweep, ms);
}

int ffvhc_cg(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countPre, int countMax, double eps,
             int* count, double* residual) {
  int prec = 8;
  int rc = prologue(c, x, A, b, bcg, &prec);
  if (rc) return rc;
  if (countPre != 1)
    return fail(-1, "the device CG implements countPreConditioner = 1 (z = omega r / Ap)");
  if (countMax < 1) {
    const std::vector<int> ids = selected_cases(c);
    for (size_t i = 0; i < ids.size(); i++)
      if (count) count[i] = 1;
    return 0;
  }
  if (prec == 8)
    rc = cg_core<double>(c, x, coef<double>(c, A), (const double*)c->fields[b].d[0], bcg, omega, countMax,
                         eps, c->dst, c->hst, nullptr, nullptr);
  else
    rc = cg_core<float>(c, x, coef<float>(c, A), (const float*)c->fields[b].d[0], bcg, omega, countMax, eps,
                        c->dst, c->hst, nullptr, nullptr);
  if (rc) return rc;
  rc = sync_state(c, c->dst, c->hst);
  if (rc) return rc;
  return report(c, c->hst, count, residual);
}

int ffvhc_cg_timing(ffvhc_ctx* c, int x, const int* A, int b, double omega, int niter, double* ms3) {
  int prec = 8;
  int rc = prologue(c, x, A, b, BC_NULL, &prec);
  if (rc) return rc;
  if (niter < 1 || !ms3) return fail(-1, "niter >= 1 and ms[3] required");
  return prec == 8 ? cg_timing_run<double>(c, x, A, b, omega, niter, ms3)
                   : cg_timing_run<float>(c, x, A, b, omega, niter, ms3);
}

int ffvhc_bicgstab_timing(ffvhc_ctx* c, int x, const int* A, int b, double omega, int niter, double* ms5) {
  int prec = 8;
  int rc = prologue(c, x, A, b, BC_NULL, &prec);
  if (rc) return rc;
  if (niter < 1 || !ms5) return fail(-1, "niter >= 1 and ms[7] required");
  return prec == 8 ? bicg_timing_run<double>(c, x, A, b, omega, niter, ms5)
                   : bicg_timing_run<float>(c, x, A, b, omega, niter, ms5);
}

int ffvhc_bicgstab_timing_mixed(ffvhc_ctx* c, int method, int x, const int* A, int b, double omega, int niter,
                                double* ms6) {
  int prec = 8;
  int rc = prologue(c, x, A, b, BC_NULL, &prec);
  if (rc) return rc;
  if (prec != 8) return fail(-1, "methods 5-9 take FP64 x, A, b");
  if (method < 5 || method > 29) return fail(-1, "method must be 5-29");
  if (niter < 1 || !ms6) return fail(-1, "niter >= 1 and ms[7] required");
  Coef7<float> M32;
  Coef7<double> M64;
  BiOpts o;
  CoarseJoin join{c};
  rc = fp64_bicg_opts(c, method == 29 ? 26 : method, A, &M32, &M64, &o);
  if (!rc) rc = coarse_join(c);
  if (rc) return rc;
  if (method == 29) {  // the inner solve of method 29: FP32 copies of A, b and x = 0
    int fa[7], fr, fe;
    for (int q = 0; q < 7; q++) {
      rc = scratch_field(c, 4, R_A + q, &fa[q]);
      if (rc) return rc;
      launch_convert(c->sm, c->dactive, c->nactive, (float*)c->fields[fa[q]].d[0],
                     (const double*)c->fields[A[q]].d[0]);
    }
    rc = scratch_field(c, 4, R_RHS, &fr);
    if (!rc) rc = scratch_field(c, 4, R_E, &fe);
    if (rc) return rc;
    launch_convert(c->sm, c->dactive, c->nactive, (float*)c->fields[fr].d[0], (const double*)c->fields[b].d[0]);
    launch_fill<float>(c->sm, c->dactive, c->nactive, (float*)c->fields[fe].d[0], (float*)c->fields[fe].h[0], 0.0f);
    return bicg_timing_run<float>(c, fe, fa, fr, omega, niter, ms6, o);
  }
  return bicg_timing_run<double>(c, x, A, b, omega, niter, ms6, o);
}

int ffvhc_ctx_set_option(ffvhc_ctx* c, int option, int value) {
  if (!c) return fail(-1, "null context");
  if (option == FFVHC_OPTION_GRAPHS) {
    c->graphs = value != 0;
    return 0;
  }
  if (option == FFVHC_OPTION_CLASSES) {
    c->classes = value != 0;
    return 0;
  }
  if (option == FFVHC_OPTION_SHADOW) {
    c->shadow_hash = value != 0;
    return 0;
  }
  if (option == FFVHC_OPTION_SETUP_REUSE) {
    c->reuse_setup = value != 0;
    c->psetup.valid = false;
    return 0;
  }
  if (option == FFVHC_OPTION_FUSE_FIN) {
    c->fuse_fin = value != 0;
    return 0;
  }
  if (option == FFVHC_OPTION_COARSE_GEMM) {
    if (value < 1) return fail(-1, "FFVHC_OPTION_COARSE_GEMM takes a block count >= 1");
    c->coarse_gemm_min = value;
    c->psetup.valid = false;
    return 0;
  }
  return fail(-1, "unknown option");
}

int ffvhc_coarse_operator(ffvhc_ctx* c, const int* A, int case_id, double* ac, double* ainv) {
  if (!c || !A || !ac || !ainv) return fail(-1, "null argument");
  for (int q = 0; q < 7; q++)
    if (!field_ok(c, A[q], 8)) return fail(-1, "A bands must be FP64 fields");
  if (case_id < 0 || case_id >= (int)c->cases.size() || !c->cases[case_id].used) return fail(-4, "no such case");
  const std::vector<int> saved = c->sel;
  int rc = ffvhc_ctx_select(c, &case_id, 1);
  if (rc) return rc;
  rc = refresh_active(c);
  BiOpts o;
  if (!rc) rc = coarse_setup(c, coef<double>(c, A), &o, ac, ainv);
  const int rc2 = ffvhc_ctx_select(c, saved.empty() ? nullptr : saved.data(), (int)saved.size());
  return rc ? rc : rc2;
}

int ffvhc_stencil_classes(ffvhc_ctx* c, const int* A, int* n_uniform, int* n_other) {
  if (!c || !A || !n_uniform || !n_other) return fail(-1, "null argument");
  for (int q = 0; q < 7; q++)
    if (!field_ok(c, A[q], 8)) return fail(-1, "A bands must be FP64 fields");
  int rc = refresh_active(c);
  if (rc) return rc;
  const bool saved = c->classes;
  c->classes = true;
  BiOpts o;
  rc = class_setup(c, coef<double>(c, A), &o);
  c->classes = saved;
  *n_uniform = o.nreg;
  *n_other = o.nirr;
  return rc;
}

int ffvhc_bicgstab(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countPre, int countMax,
                   double eps, int* count, double* residual) {
  int prec = 8;
  int rc = prologue(c, x, A, b, bcg, &prec);
  if (rc) return rc;
  if (countPre != 1)
    return fail(-1, "the device BiCGSTAB implements countPreConditioner = 1 (M v = omega v / Ap)");
  if (countMax < 1) {
    const std::vector<int> ids = selected_cases(c);
    for (size_t i = 0; i < ids.size(); i++)
      if (count) count[i] = 1;
    return 0;
  }
  c->bicg_round = 0;
  c->bicg_hint_key = (prec == 8 ? -100 : -200) - bcg;  // per BC group: the P and U solves of a step differ
  if (prec == 8)
    rc = bicgstab_core<double>(c, x, coef<double>(c, A), (const double*)c->fields[b].d[0], bcg, omega,
                               countMax, eps, c->dst, c->hst, nullptr, nullptr);
  else
    rc = bicgstab_core<float>(c, x, coef<float>(c, A), (const float*)c->fields[b].d[0], bcg, omega,
                              countMax, eps, c->dst, c->hst, nullptr, nullptr);
  c->bicg_round = -1;
  if (rc) return rc;
  rc = sync_state(c, c->dst, c->hst);
  if (rc) return rc;
  return report(c, c->hst, count, residual);
}

// FP32 inner solve in FP64 defect correction.  count = inner iterations summed over the
// corrections, residual = final FP64 relative residual, outer = corrections applied.
int ffvhc_solve_mixed(ffvhc_ctx* c, int method, int x, const int* A, int b, int bcg, double omega, int countMax,
                      double eps, double inner_eps, int* count, double* residual, int* outer) {
  int prec = 8;
  int rc = prologue(c, x, A, b, bcg, &prec);
  if (rc) return rc;
  if (prec != 8) return fail(-1, "mixed precision takes FP64 x, A, b");
  CoarseJoin join{c};
  if (method < 0 || method > 29)
    return fail(-1, "method: 0 RBGS, 1 Jacobi-PCG, 2 Jacobi-BiCGSTAB, 3 Jacobi-BiCGSTAB with reliable updates, "
                    "4 the same with a seeded shadow residual, 5 FP64 BiCGSTAB with an FP32 red-black GS "
                    "preconditioner, 6 the same all in FP64 (reference point), 7 FP64 Jacobi-BiCGSTAB with an FP32 "
                    "matrix, 8 as 7 with the block red-black GS preconditioner (FP64 arithmetic), 9 as 8 with the "
                    "symmetric red-black-red sweep, 10 as 8 with the seeded hash vector as shadow residual, "
                    "11-16 as 8 with 1-6 red-black SOR sweeps on a compact FP32/FP16 copy of the matrix, "
                    "17-22 as 11-16 plus a coarse (one value per block) correction, 23-28 the same correction "
                    "applied before the sweeps (multiplicative), 29 method 26's BiCGSTAB on FP32 vectors inside "
                    "FP64 defect correction");
  if (method == 6) {
    Coef7<float> M32;
    Coef7<double> M64;
    BiOpts o;
    rc = fp64_bicg_opts(c, method, A, &M32, &M64, &o);
    if (rc) return rc;
    rc = bicgstab_core<double>(c, x, M64, (const double*)c->fields[b].d[0], bcg, omega, countMax, eps, c->dst,
                               c->hst, nullptr, nullptr, o);
    if (rc) return rc;
    rc = sync_state(c, c->dst, c->hst);
    if (rc) return rc;
    const std::vector<int> ids6 = selected_cases(c);
    for (size_t i = 0; i < ids6.size(); i++)
      if (outer) outer[i] = 0;
    return report(c, c->hst, count, residual);
  }
  if (method == 29) return two_level_defect_correction(c, x, A, b, bcg, omega, countMax, eps, inner_eps, count,
                                                       residual, outer);
  if (method == 5 || method >= 7) {  // FP64 recurrence, dots and residual; FP32 matrix
    Coef7<float> M32;
    Coef7<double> K64;
    BiOpts o;
    if (c->step_solve && method >= 11 && psetup_match(c, method, A)) {
      // the time step's pressure matrix is the one this setup was made from
      K64 = coef<double>(c, A);
      int fa[7];
      for (int q = 0; q < 7; q++) {
        rc = scratch_field(c, 4, R_A + q, &fa[q]);
        if (rc) return rc;
      }
      M32 = coef<float>(c, fa);
      o = c->psetup.o;
      o.A32 = &M32;
    } else {
      rc = fp64_bicg_opts(c, method, A, &M32, &K64, &o);
      if (rc) return rc;
      if (c->step_solve && method >= 11) psetup_store(c, method, A, o);
    }
    // The Krylov recurrence converges for the FP32-stored matrix.  Its answer is accepted only on the
    // FP64 residual ||b - A x|| of the FP64 matrix; a case above eps (a matrix whose coefficients are
    // not all FP32-representable) restarts from that residual, so the reported residual is always
    // the FP64 one.
    int fe;
    rc = scratch_field(c, 4, R_E, &fe);
    if (rc) return rc;
    const double* b64 = (const double*)c->fields[b].d[0];
    const std::vector<int> ids5 = selected_cases(c);
    std::vector<int> outers(c->cases.size(), -1);
    launch_solve_init(c->sm, c->dst2, (int)c->cases.size(), omega, eps, 100000);
    int o5 = 0;
    bool done = false;
    auto defect = [&]() -> int {
      const FieldRec X = c->fields[x];
      launch_defect(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst2, K64, b64, (const double*)X.d[0],
                    (const double*)X.h[0], (float*)c->fields[fe].d[0], c->dpart, (size_t)c->cap, o5);
      fin(c, c->dst2, FIN_DEFECT, o5);
      HIPCHK(hipGetLastError());
      int r = read_state(c, c->dst2, c->hst2, ids5, &done);
      for (int id : ids5)
        if (c->hst2[id].done && outers[id] < 0) outers[id] = o5;
      return r;
    };
    rc = defect();
    if (rc) return rc;
    std::vector<int> its(c->cases.size(), 0);
    // another round while a case is above eps and has budget left (a case's rounds depend on it alone)
    auto more = [&]() {
      for (int id : ids5)
        if (!c->hst2[id].done && its[id] < countMax) return true;
      return false;
    };
    while (more()) {
      // each case's budget: countMax less its own iterations so far (dst2 used)
      c->bicg_round = o5;
      c->bicg_hint_key = method;
      rc = bicgstab_core<double>(c, x, K64, b64, bcg, omega, countMax, eps, c->dst, c->hst, c->dst2, nullptr, o);
      c->bicg_round = -1;
      if (rc) return rc;
      launch_add_used(c->sm, c->dst2, c->dst, (int)c->cases.size());
      bool inner_done;
      rc = read_state(c, c->dst, c->hst, ids5, &inner_done);
      if (rc) return rc;
      for (int id : ids5) its[id] += c->hst[id].count;
      o5++;
      rc = defect();
      if (rc) return rc;
    }
    int status = 0;
    for (size_t i = 0; i < ids5.size(); i++) {
      const SolveState& s = c->hst2[ids5[i]];
      if (count) count[i] = its[ids5[i]];
      if (residual) residual[i] = s.residual;
      if (outer) outer[i] = outers[ids5[i]] < 0 ? o5 : outers[ids5[i]];
      if (s.status) status = 1;
    }
    if (status) set_error("residual is NaN or Inf");
    return status;
  }
  if (method >= 3) {
    if (!(inner_eps > 0.0 && inner_eps < 1.0)) return fail(-1, "reliable-update factor must be in (0, 1)");
    return bicgstab_reliable(c, x, A, b, bcg, omega, countMax, eps, inner_eps, count, residual, outer, method == 4);
  }
  const std::vector<int> ids = selected_cases(c);
  int fa[7], fr, fe;
  for (int q = 0; q < 7; q++) {
    rc = scratch_field(c, 4, R_A + q, &fa[q]);
    if (rc) return rc;
  }
  rc = scratch_field(c, 4, R_RHS, &fr);
  if (!rc) rc = scratch_field(c, 4, R_E, &fe);
  if (rc) return rc;
  if (method == 0) {
    rc = ensure_second_buffer(c, fe);
    if (rc) return rc;
  }
  // FP32 copy of the matrix
  c->setup_serial++;
  for (int q = 0; q < 7; q++)
    launch_convert(c->sm, c->dactive, c->nactive, (float*)c->fields[fa[q]].d[0],
                   (const double*)c->fields[A[q]].d[0]);
  Coef7<double> K64 = coef<double>(c, A);
  Coef7<float> K32 = coef<float>(c, fa);
  const FieldRec X = c->fields[x], E = c->fields[fe], Rf = c->fields[fr];
  std::vector<int> inner_total(c->cases.size(), 0), outers(c->cases.size(), -1);
  const int max_outer = 64;
  launch_solve_init(c->sm, c->dst2, (int)c->cases.size(), omega, eps, max_outer);
  for (int o = 0;; o++) {
    launch_defect(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst2, K64, (const double*)c->fields[b].d[0],
                  (const double*)X.d[0], (const double*)X.h[0], (float*)Rf.d[0], c->dpart, (size_t)c->cap, o);
    fin(c, c->dst2, FIN_DEFECT, o);
    HIPCHK(hipGetLastError());
    bool done = false;
    rc = read_state(c, c->dst2, c->hst2, ids, &done);
    if (rc) return rc;
    for (int id : ids)
      if (c->hst2[id].done && outers[id] < 0) outers[id] = o;
    if (done || o >= max_outer) break;
    launch_fill<float>(c->sm, c->dactive, c->nactive, (float*)E.d[0], (float*)E.h[0], 0.0f);
    if (method == 0)
      rc = rbgs_core<float>(c, fe, K32, (const float*)Rf.d[0], BC_HOM, omega, countMax, inner_eps, c->dst, c->hst,
                            c->dst2, nullptr);
    else if (method == 1)
      rc = cg_core<float>(c, fe, K32, (const float*)Rf.d[0], BC_HOM, omega, countMax, inner_eps, c->dst, c->hst,
                          c->dst2, nullptr);
    else
      rc = bicgstab_core<float>(c, fe, K32, (const float*)Rf.d[0], BC_HOM, omega, countMax, inner_eps, c->dst,
                                c->hst, c->dst2, nullptr);
    if (rc) return rc;
    rc = sync_state(c, c->dst, c->hst);
    if (rc) return rc;
    for (int id : ids)
      if (!c->hst2[id].done) inner_total[id] += std::max(c->hst[id].count, 0);
    launch_add_correction(c->sm, c->dactive, c->nactive, c->dblk, c->dst2, c->dst, (double*)X.d[0],
                          (const float*)E.d[0]);
    launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const double*)X.d[0],
                              (double*)X.h[0], bcg, 1, c->multilevel);
    HIPCHK(hipGetLastError());
  }
  rc = sync_state(c, c->dst2, c->hst2);
  if (rc) return rc;
  int status = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const SolveState& s = c->hst2[ids[i]];
    if (count) count[i] = inner_total[ids[i]];
    if (residual) residual[i] = s.residual;
    if (outer) outer[i] = outers[ids[i]] < 0 ? max_outer : outers[ids[i]];
    if (s.status) status = 1;
  }
  if (status) set_error("residual is NaN or Inf");
  return status;
}

int ffvhc_residual(ffvhc_ctx* c, int x, const int* A, int b, double* residual) {
  int prec = 8;
  int rc = prologue(c, x, A, b, BC_P, &prec);
  if (rc) return rc;
  if (prec != 8) return fail(-1, "residual takes FP64 x, A, b");
  int fr;
  rc = scratch_field(c, 4, R_RHS, &fr);
  if (rc) return rc;
  const std::vector<int> ids = selected_cases(c);
  launch_solve_init(c->sm, c->dst2, (int)c->cases.size(), 1.0, -1.0, 1);
  launch_defect(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst2, coef<double>(c, A),
                (const double*)c->fields[b].d[0], (const double*)c->fields[x].d[0],
                (const double*)c->fields[x].h[0], (float*)c->fields[fr].d[0], c->dpart, (size_t)c->cap, 0);
  fin(c, c->dst2, FIN_DEFECT, 0);
  HIPCHK(hipGetLastError());
  rc = sync_state(c, c->dst2, c->hst2);
  if (rc) return rc;
  for (size_t i = 0; i < ids.size(); i++)
    if (residual) residual[i] = c->hst2[ids[i]].residual;
  return 0;
}

}  // extern "C"

// Per-case force triples into f[3 * selected case] (0 for a case without blocks); `launch`
// writes the triples of the selected cases with blocks, in selection order, to its device array.
template <class Launch>
static int case_triples(ffvhc_ctx* c, double* f, Launch launch) {
  int rc = refresh_active(c);
  if (rc) return rc;
  double* dout = nullptr;
  HIPCHK(hipMallocAsync((void**)&dout, sizeof(double) * 3 * std::max(c->ncaseact, 1), c->sm));
  launch(dout);
  HIPCHK(hipGetLastError());
  std::vector<double> h(3 * std::max(c->ncaseact, 1), 0.0);
  HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * 3 * c->ncaseact, hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipFreeAsync(dout, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  // caseact lists the selected cases with blocks in selection order
  const std::vector<int> ids = selected_cases(c);
  size_t q = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    if (c->cases[ids[i]].nblk > 0) {
      for (int d = 0; d < 3; d++) f[3 * i + d] = h[3 * q + d];
      q++;
    } else {
      for (int d = 0; d < 3; d++) f[3 * i + d] = 0.0;
    }
  }
  return 0;
}

extern "C" {

int ffvhc_force_p(ffvhc_ctx* c, int p, int cut_id, double* f) {
  if (!c || !f) return fail(-1, "null argument");
  if (!field_ok(c, p, 8)) return fail(-1, "p must be an FP64 field");
  if (cut_id < 1 || cut_id > 255) return fail(-1, "cut_id must be in 1..255");
  return case_triples(c, f, [&](double* dout) {
    launch_force_p(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase,
                   (const double*)c->fields[p].d[0], (const double*)c->fields[p].h[0], c->dcid, c->dflag,
                   (size_t)c->cap * NC, cut_id, c->dpart, (size_t)c->cap, dout);
  });
}

int ffvhc_force_p_noalloc(ffvhc_ctx* c, int p, int cut_id, double* f) {
  if (!c || !f) return fail(-1, "null argument");
  if (!field_ok(c, p, 8)) return fail(-1, "p must be an FP64 field");
  if (cut_id < 1 || cut_id > 255) return fail(-1, "cut_id must be in 1..255");
  if (c->cap > FFVHC_NOALLOC_SLOTS)
    return fail(-1, "ffvhc_force_p_noalloc: the context has " + std::to_string(c->cap) +
                        " block slots, the scratch holds " + std::to_string(FFVHC_NOALLOC_SLOTS));
  if (!c->dscratch) return fail(-2, "no scratch address for this device");
  int rc = refresh_active(c);
  if (rc) return rc;
  // scratch: block partials [3][FFVHC_NOALLOC_SLOTS] (force_p_kernel writes every active slot
  // before force_sum_kernel reads a case's slots, all of which are active), then the case triples
  // [MAX_CASES][3] (written for every active case); nothing needs clearing
  double* part = c->dscratch;
  double* out = c->dscratch + 3 * (size_t)FFVHC_NOALLOC_SLOTS;
  launch_force_p(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase,
                 (const double*)c->fields[p].d[0], (const double*)c->fields[p].h[0], c->dcid, c->dflag,
                 (size_t)c->cap * NC, cut_id, part, (size_t)FFVHC_NOALLOC_SLOTS, out);
  HIPCHK(hipGetLastError());
  double h[3 * MAX_CASES];
  if (c->ncaseact)
    HIPCHK(hipMemcpyAsync(h, out, sizeof(double) * 3 * c->ncaseact, hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  // f in selection order, zeros for a selected case without blocks (as ffvhc_force_p)
  size_t i = 0, q = 0;
  auto put = [&](int id) {
    const bool has = c->cases[id].nblk > 0;
    for (int d = 0; d < 3; d++) f[3 * i + d] = has ? h[3 * q + d] : 0.0;
    q += has;
    i++;
  };
  if (!c->sel.empty()) {
    for (int id : c->sel)
      if (id >= 0 && id < (int)c->cases.size() && c->cases[id].used) put(id);
  } else {
    for (int id = 0; id < (int)c->cases.size(); id++)
      if (c->cases[id].used) put(id);
  }
  return 0;
}

int ffvhc_force_v(ffvhc_ctx* c, int ux, int uy, int uz, int cut_id, double* f) {
  if (!c || !f) return fail(-1, "null argument");
  if (!field_ok(c, ux, 8) || !field_ok(c, uy, 8) || !field_ok(c, uz, 8))
    return fail(-1, "ux, uy, uz must be FP64 fields");
  if (cut_id < 1 || cut_id > 255) return fail(-1, "cut_id must be in 1..255");
  const double* u[3] = {(const double*)c->fields[ux].d[0], (const double*)c->fields[uy].d[0],
                        (const double*)c->fields[uz].d[0]};
  const double* hu[3] = {(const double*)c->fields[ux].h[0], (const double*)c->fields[uy].h[0],
                         (const double*)c->fields[uz].h[0]};
  return case_triples(c, f, [&](double* dout) {
    launch_force_v(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase, u, hu, c->dcut,
                   c->dcid, c->dflag, (size_t)c->cap * NC, cut_id, c->dpart, (size_t)c->cap, dout);
  });
}

int ffvhc_field_stats(ffvhc_ctx* c, int f, double* out) {
  if (!c || !out) return fail(-1, "null argument");
  if (!field_ok(c, f, 8)) return fail(-1, "stats take an FP64 field");
  int rc = refresh_active(c);
  if (rc) return rc;
  const std::vector<int> ids = selected_cases(c);
  const int n = std::max(c->ncaseact, 1);
  double* dout = nullptr;
  HIPCHK(hipMallocAsync((void**)&dout, sizeof(double) * 5 * n, c->sm));
  launch_stats(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dcase, (const double*)c->fields[f].d[0],
               c->dpart, (size_t)c->cap, dout);
  HIPCHK(hipGetLastError());
  std::vector<double> h(5 * n, 0.0);
  HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * 5 * c->ncaseact, hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipFreeAsync(dout, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  size_t q = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const bool has = c->cases[ids[i]].nblk > 0;
    for (int d = 0; d < 5; d++) out[5 * i + d] = has ? h[5 * q + d] : 0.0;
    if (has) q++;
  }
  return 0;
}

int ffvhc_cfl_check(ffvhc_ctx* c, int ux, double* cfl) {
  std::vector<double> st;
  const std::vector<int> ids = c ? selected_cases(c) : std::vector<int>();
  st.resize(5 * std::max<size_t>(ids.size(), 1));
  int rc = ffvhc_field_stats(c, ux, st.data());
  if (rc) return rc;
  int status = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const CaseRec& r = c->cases[ids[i]];
    // Solver::PrintStats (Solver.cpp:2655-2676): dx = RootBlockLength / 2^maxLevel
    const double dx = r.root_length / (double)(1 << r.max_level);
    const double v = st[5 * i + 3] * r.params.dt / dx;
    if (cfl) cfl[i] = v;
    if (v > 0.9) status = 2;
  }
  if (status) set_error("CFL guard: |ux|max dt / dx > 0.9");
  return status;
}

int ffvhc_case_set_step(ffvhc_ctx* c, int id, int step, int accel_steps) {
  if (!c || id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  if (step < 0 || accel_steps < 0) return fail(-1, "step and accel_steps must be >= 0");
  c->cases[id].step = step;
  c->cases[id].accel_steps = accel_steps;
  return upload_case_meta(c, id);
}

int ffvhc_case_step(const ffvhc_ctx* c, int id) {
  if (!c || id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  return c->cases[id].step;
}

int ffvhc_step(ffvhc_ctx* c, ffvhc_step_fields* F, const ffvhc_step_params* P, ffvhc_step_info* info,
               double* stage_ms) {
  if (!c || !F || !P) return fail(-1, "null argument");
  const int req[] = {F->ux, F->uy, F->uz, F->p, F->uc[0], F->uc[1], F->uc[2], F->ucp[0], F->ucp[1], F->ucp[2],
                     F->ud[0], F->ud[1], F->ud[2], F->udp[0], F->udp[1], F->udp[2], F->v[0], F->v[1], F->v[2],
                     F->v[3], F->v[4], F->v[5], F->A[0], F->A[1], F->A[2], F->A[3], F->A[4], F->A[5], F->A[6],
                     F->b, F->lapp};
  for (int f : req)
    if (!field_ok(c, f, 8)) return fail(-1, "time-step fields must be allocated FP64 fields");
  const int opt[] = {F->t, F->f[0], F->f[1], F->f[2]};
  for (int f : opt)
    if (f != -1 && !field_ok(c, f, 8)) return fail(-1, "t / f must be -1 or FP64 fields");
  if (P->solver < 0 || P->solver > 3)
    return fail(-1, "solver: 0 RBGS, 1 RBGS_WithFixedAverage, 2 Jacobi-PCG, 3 Jacobi-BiCGSTAB");
  if (P->solver >= 2 && P->countPre != 1) return fail(-1, "countPre must be 1");
  if (P->mixed && P->mixed_method < 0 && P->solver == 1) return fail(-1, "mixed precision takes RBGS, PCG or BiCGSTAB");
  if (P->u_implicit && P->countMax_u < 0) return fail(-1, "countMax_u must be >= 0");
  int rc = refresh_active(c);
  if (rc) return rc;
  const std::vector<int> ids = selected_cases(c);
  for (int id : ids) {
    rc = upload_case_meta(c, id);
    if (rc) return rc;
  }
  auto D = [&](int f) { return (double*)c->fields[f].d[0]; };
  auto H = [&](int f) { return (double*)c->fields[f].h[0]; };
  StepFields S;
  const int u[3] = {F->ux, F->uy, F->uz};
  for (int d = 0; d < 3; d++) {
    S.u[d] = D(u[d]);
    S.hu[d] = H(u[d]);
    S.uc[d] = D(F->uc[d]);
    S.ud[d] = D(F->ud[d]);
    S.ucp[d] = D(F->ucp[d]);
    S.udp[d] = D(F->udp[d]);
    S.f[d] = F->f[d] < 0 ? nullptr : D(F->f[d]);
  }
  S.p = D(F->p);
  S.hp = H(F->p);
  for (int q = 0; q < 6; q++) S.v[q] = D(F->v[q]);
  for (int q = 0; q < 7; q++) S.A[q] = D(F->A[q]);
  S.b = D(F->b);
  S.lapp = D(F->lapp);
  S.t = F->t < 0 ? nullptr : D(F->t);
  const size_t capnc = (size_t)c->cap * NC;
  auto fill_u = [&]() {
    for (int d = 0; d < 3; d++)
      launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S.u[d], S.hu[d], BC_UX + d, 5,
                                c->multilevel);
  };
  hipEvent_t ev[4];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], c->sm));
  const size_t nc = ids.size();
  std::vector<int> cntu(3 * nc, 0);
  std::vector<double> resu(3 * nc, 0.0);
  if (P->u_implicit) {
    // UpdateUX / UY / UZ: matrix and RHS, the component's matrix BC, BiCGSTAB, ghost fill
    for (int d = 0; d < 3; d++) {
      launch_abd_u(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, c->dcut, c->dflag, capnc, d, P->alpha_u);
      launch_matrix_bc(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, BC_UX + d);
      HIPCHK(hipGetLastError());
      int stu = ffvhc_bicgstab(c, u[d], F->A, F->b, BC_UX + d, P->omega_u, 1, P->countMax_u, P->eps_u,
                               cntu.data() + d * nc, resu.data() + d * nc);
      if (stu < 0) return stu;
      launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S.u[d], S.hu[d], BC_UX + d, 5,
                                c->multilevel);
    }
  } else {
    // UpdateUXe / UYe / UZe: the three components are independent (face velocities are fixed)
    launch_predict(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, c->dcut, c->dflag, capnc);
    fill_u();
  }
  HIPCHK(hipEventRecord(ev[1], c->sm));
  // UpdateP
  launch_p1(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, c->dcut, c->dflag, capnc);
  fill_u();
  launch_ab_p(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, c->dcut, c->dflag, capnc);
  if (P->ref_active) {
    std::vector<int> cand;
    for (int id : ids) {
      const CaseRec& r = c->cases[id];
      for (int q = 0; q < r.nblk; q++) {
        const DevBlock& B = c->hblk[r.slot_begin + q];
        const double L = B.dx * NB;
        bool inside = true;
        for (int a = 0; a < 3; a++)
          if (!(B.org[a] <= P->ref[a] && P->ref[a] <= B.org[a] + L)) inside = false;
        if (inside) cand.push_back(r.slot_begin + q);
      }
    }
    if (!cand.empty()) {
      int* dc = nullptr;
      HIPCHK(hipMallocAsync((void**)&dc, sizeof(int) * cand.size(), c->sm));
      HIPCHK(hipMemcpyAsync(dc, cand.data(), sizeof(int) * cand.size(), hipMemcpyHostToDevice, c->sm));
      launch_refval(c->sm, dc, (int)cand.size(), c->dblk, S, P->ref[0], P->ref[1], P->ref[2], P->ref_p);
      HIPCHK(hipFreeAsync(dc, c->sm));
    }
  }
  launch_matrix_bc(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, BC_P);
  HIPCHK(hipGetLastError());
  std::vector<int> cnt(nc, 0), outer(nc, 0);
  std::vector<double> res(nc, 0.0);
  int st = 0;
  if (P->mixed) {
    const int method = P->mixed_method >= 0 ? P->mixed_method : P->solver == 0 ? 0 : (P->solver == 2 ? 1 : 2);
    c->step_solve = true;  // A is this step's matrix, made from the geometry and the BC types
    c->step_ref_active = P->ref_active != 0;
    for (int a = 0; a < 3; a++) c->step_ref[a] = P->ref_active ? P->ref[a] : 0.0;
    st = ffvhc_solve_mixed(c, method, F->p, F->A, F->b, BC_P, P->omega, P->countMax, P->eps, P->inner_eps,
                           cnt.data(), res.data(), outer.data());
    c->step_solve = false;
  } else if (P->solver == 0) {
    st = ffvhc_rbgs(c, F->p, F->A, F->b, BC_P, P->omega, P->countMax, P->eps, cnt.data(), res.data());
  } else if (P->solver == 1) {
    st = ffvhc_rbgs_fixed_average(c, F->p, F->A, F->b, BC_P, P->omega, P->countMax, P->eps, cnt.data(),
                                  res.data());
  } else if (P->solver == 2) {
    st = ffvhc_cg(c, F->p, F->A, F->b, BC_P, P->omega, P->countPre, P->countMax, P->eps, cnt.data(), res.data());
  } else {
    st = ffvhc_bicgstab(c, F->p, F->A, F->b, BC_P, P->omega, P->countPre, P->countMax, P->eps, cnt.data(),
                        res.data());
  }
  if (st < 0) return st;
  launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S.p, S.hp, BC_P, 1, c->multilevel);
  HIPCHK(hipEventRecord(ev[2], c->sm));
  // UpdateU
  launch_corr(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S, c->dcut, c->dflag, capnc);
  fill_u();
  launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S.p, S.hp, BC_P, 1, c->multilevel);
  launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, S.lapp, H(F->lapp), BC_NULL, 1,
                            c->multilevel);
  HIPCHK(hipEventRecord(ev[3], c->sm));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(ev[3]));
  float t01 = 0.f, t12 = 0.f, t23 = 0.f;
  HIPCHK(hipEventElapsedTime(&t01, ev[0], ev[1]));
  HIPCHK(hipEventElapsedTime(&t12, ev[1], ev[2]));
  HIPCHK(hipEventElapsedTime(&t23, ev[2], ev[3]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (stage_ms) {
    stage_ms[0] = t01 + t12 + t23;
    stage_ms[1] = 0.0;
    stage_ms[2] = t01;
    stage_ms[3] = 0.0;
    stage_ms[4] = 0.0;
    stage_ms[5] = t12;
    stage_ms[6] = t23;
  }
  for (int id : ids) c->cases[id].step++;
  for (int d = 0; d < 3; d++) {
    std::swap(F->uc[d], F->ucp[d]);
    std::swap(F->ud[d], F->udp[d]);
  }
  for (size_t i = 0; i < nc; i++)
    if (info) {
      info[i].countP = cnt[i];
      info[i].residualP = res[i];
      info[i].outer = outer[i];
      info[i].status = (std::isnan(res[i]) || std::isinf(res[i])) ? 1 : 0;
      for (int d = 0; d < 3; d++) {
        info[i].countU[d] = cntu[d * nc + i];
        info[i].residualU[d] = resu[d * nc + i];
        if (std::isnan(resu[d * nc + i]) || std::isinf(resu[d * nc + i])) info[i].status = 1;
      }
    }
  if (st) set_error("Poisson residual is NaN or Inf");
  return st;
}

}  // extern "C"
