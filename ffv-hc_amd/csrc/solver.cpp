// Linear solvers behind the C ABI (FFVILS.h): fused device-resident RBGS and Jacobi-PCG, and the
// FP32-inner / FP64-defect-correction mixed-precision drivers.
#include <algorithm>
#include <cstring>

#include "kernels.h"
#include "runtime.h"

using namespace ffvhc;

namespace {

// Iterations launched between host checks of the per-case `done` flags; the device skips the
// launches of finished cases, so a batch only bounds how long the host waits.
constexpr int kBatchMin = 8;
constexpr int kBatchMax = 256;

// scratch roles
enum { R_R = 1, R_P = 2, R_Q = 3, R_E = 4, R_RHS = 5, R_A = 10, R_BI = 20 };

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
inline void fin(ffvhc_ctx* c, SolveState* st, int op, int arg) {
  launch_finalize(c->sm, c->dcaseact, c->ncaseact, c->dcase, st, c->dpart, (size_t)c->cap, op, arg);
}

// RBGS on field x (ping-pong), matrix K, rhs bv; state in st. Result left in buffer 0.
template <class T>
int rbgs_core(ffvhc_ctx* c, int xf, Coef7<T> K, const T* bv, int bcg, double omega, int countMax, double eps,
              SolveState* st, std::vector<SolveState>& hst, const SolveState* outer, int* launched) {
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
                          hs[(n - 1) & 1], hs[n & 1], c->dpart, (size_t)c->cap, n, bcg);
      fin(c, st, FIN_RBGS, n);
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
                  double eps, SolveState* st, std::vector<SolveState>& hst, const SolveState* outer, int* launched) {
  const std::vector<int> ids = selected_cases(c);
  BiVecs<T> V;
  int rc0 = bicg_vectors<T>(c, &V);
  if (rc0) return rc0;
  const FieldRec X = c->fields[xf];  // copy taken after the scratch allocations
  launch_solve_init(c->sm, st, (int)c->cases.size(), omega, eps, countMax, outer);
  launch_bicgstab_init<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, st, K, bv, (const T*)X.d[0],
                          (const T*)X.h[0], V, c->dpart, (size_t)c->cap);
  fin(c, st, FIN_BI_INIT, 0);
  int n = 1, batch = kBatchMin;
  bool done = false;
  while (!done) {
    int last = std::min(n + batch - 1, countMax);
    for (; n <= last; n++)
      launch_bicgstab_iter<T>(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase, st, K,
                              (T*)X.d[0], V, c->dpart, (size_t)c->cap, n);
    HIPCHK(hipGetLastError());
    int rc = read_state(c, st, hst, ids, &done);
    if (rc) return rc;
    if (n > countMax) done = true;
    batch = std::min(batch * 2, kBatchMax);
  }
  if (launched) *launched = n - 1;
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
  hipEvent_t e0, e1;
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
int bicg_timing_run(ffvhc_ctx* c, int x, const int* A, int b, double omega, int niter, double* ms) {
  BiVecs<T> V;
  int rc = bicg_vectors<T>(c, &V);
  if (rc) return rc;
  const FieldRec X = c->fields[x];
  Coef7<T> K = coef<T>(c, A);
  launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, -1.0, niter);
  launch_bicgstab_init<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K,
                          (const T*)c->fields[b].d[0], (const T*)X.d[0], (const T*)X.h[0], V, c->dpart,
                          (size_t)c->cap);
  fin(c, c->dst, FIN_BI_INIT, 0);
  // stages in launch order: P, Q, fin(alpha), S, T, fin(gamma), XR, fin(convergence, beta)
  static const int bit_of[8] = {1, 2, 0, 4, 8, 0, 16, 0};
  static const int fin_of[8] = {-1, -1, FIN_BI_ALPHA, -1, -1, FIN_BI_GAMMA, -1, FIN_BI_XR};
  static const int slot_of[8] = {0, 1, 5, 2, 3, 5, 4, 5};
  std::vector<hipEvent_t> ev(9 * niter);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int n = 1; n <= niter; n++) {
    hipEvent_t* e = &ev[9 * (n - 1)];
    HIPCHK(hipEventRecord(e[0], c->sm));
    for (int k = 0; k < 8; k++) {
      if (bit_of[k])
        launch_bicgstab_iter<T>(c->sm, c->dactive, c->nactive, c->dcaseact, c->ncaseact, c->dblk, c->dcase, c->dst,
                                K, (T*)X.d[0], V, c->dpart, (size_t)c->cap, n, bit_of[k] | 32);
      else
        fin(c, c->dst, fin_of[k], n);
      HIPCHK(hipEventRecord(e[k + 1], c->sm));
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->sm));
  for (int k = 0; k < 6; k++) ms[k] = 0.0;
  for (int n = 0; n < niter; n++)
    for (int k = 0; k < 8; k++) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev[9 * n + k], ev[9 * n + k + 1]));
      ms[slot_of[k]] += t;
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return 0;
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

int ffvhc_rbgs(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax, double eps, int* count,
               double* residual) {
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
                           eps, c->dst, c->hst, nullptr, nullptr);
  else
    rc = rbgs_core<float>(c, x, coef<float>(c, A), (const float*)c->fields[b].d[0], bcg, omega, countMax,
                          eps, c->dst, c->hst, nullptr, nullptr);
  if (rc) return rc;
  rc = sync_state(c, c->dst, c->hst);
  if (rc) return rc;
  return report(c, c->hst, count, residual);
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
    return fail(-1, "the device CG implements countPreConditioner = 1 (z = omega r / Ap); see DESIGN.md");
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
  if (niter < 1 || !ms5) return fail(-1, "niter >= 1 and ms[5] required");
  return prec == 8 ? bicg_timing_run<double>(c, x, A, b, omega, niter, ms5)
                   : bicg_timing_run<float>(c, x, A, b, omega, niter, ms5);
}

int ffvhc_bicgstab(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countPre, int countMax,
                   double eps, int* count, double* residual) {
  int prec = 8;
  int rc = prologue(c, x, A, b, bcg, &prec);
  if (rc) return rc;
  if (countPre != 1)
    return fail(-1, "the device BiCGSTAB implements countPreConditioner = 1 (M v = omega v / Ap); see DESIGN.md");
  if (countMax < 1) {
    const std::vector<int> ids = selected_cases(c);
    for (size_t i = 0; i < ids.size(); i++)
      if (count) count[i] = 1;
    return 0;
  }
  if (prec == 8)
    rc = bicgstab_core<double>(c, x, coef<double>(c, A), (const double*)c->fields[b].d[0], bcg, omega,
                               countMax, eps, c->dst, c->hst, nullptr, nullptr);
  else
    rc = bicgstab_core<float>(c, x, coef<float>(c, A), (const float*)c->fields[b].d[0], bcg, omega,
                              countMax, eps, c->dst, c->hst, nullptr, nullptr);
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
  if (method < 0 || method > 2) return fail(-1, "method: 0 RBGS, 1 Jacobi-PCG, 2 Jacobi-BiCGSTAB");
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
