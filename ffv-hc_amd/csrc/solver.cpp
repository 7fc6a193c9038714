// Linear solvers behind the C ABI (FFVILS.h): RBGS with the fused device-resident iteration.
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

template <class T>
Coef7<T> coef(ffvhc_ctx* c, const int* A) {
  Coef7<T> k;
  for (int q = 0; q < 7; q++) k.p[q] = (const T*)c->fields[A[q]].d[0];
  return k;
}

int read_state(ffvhc_ctx* c, const std::vector<int>& ids, bool* all_done) {
  HIPCHK(hipMemcpyAsync(c->hst.data(), c->dst, sizeof(SolveState) * c->cases.size(), hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  *all_done = true;
  for (int id : ids)
    if (!c->hst[id].done) *all_done = false;
  return 0;
}

int check_args(ffvhc_ctx* c, int x, const int* A, int b, int prec) {
  if (!c || !A) return fail(-1, "null argument");
  if (!field_ok(c, x, prec) || !field_ok(c, b, prec)) return fail(-1, "x/b must be fields of one precision");
  for (int q = 0; q < 7; q++)
    if (!field_ok(c, A[q], prec)) return fail(-1, "A bands must be fields of the same precision");
  return 0;
}

template <class T>
int rbgs_run(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int countMax, double eps) {
  int rc = ensure_second_buffer(c, x);
  if (rc) return rc;
  FieldRec& X = c->fields[x];
  const std::vector<int> ids = selected_cases(c);
  launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, eps, countMax);
  Coef7<T> K = coef<T>(c, A);
  const T* bv = (const T*)c->fields[b].d[0];
  T* xs[2] = {(T*)X.d[0], (T*)X.d[1]};
  T* hs[2] = {(T*)X.h[0], (T*)X.h[1]};
  int n = 1, batch = kBatchMin;
  bool done = false;
  while (!done) {
    int last = std::min(n + batch - 1, countMax + 1);
    for (; n <= last; n++)
      launch_rbgs_iter<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, bv, xs[(n - 1) & 1], xs[n & 1],
                          hs[(n - 1) & 1], hs[n & 1], c->dpart, (size_t)c->cap, n, bcg);
    HIPCHK(hipGetLastError());
    rc = read_state(c, ids, &done);
    if (rc) return rc;
    if (n > countMax + 1) done = true;
    batch = std::min(batch * 2, kBatchMax);
  }
  launch_select<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dst, xs[0], xs[1], hs[0], hs[1]);
  HIPCHK(hipGetLastError());
  return 0;
}

template <class T>
int rbgs_fixed_run(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int nsweep, double* ms) {
  int rc = ensure_second_buffer(c, x);
  if (rc) return rc;
  FieldRec& X = c->fields[x];
  launch_solve_init(c->sm, c->dst, (int)c->cases.size(), omega, -1.0, nsweep);
  Coef7<T> K = coef<T>(c, A);
  const T* bv = (const T*)c->fields[b].d[0];
  T* xs[2] = {(T*)X.d[0], (T*)X.d[1]};
  T* hs[2] = {(T*)X.h[0], (T*)X.h[1]};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->sm));
  for (int n = 1; n <= nsweep; n++)
    launch_rbgs_iter<T>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, c->dst, K, bv, xs[(n - 1) & 1], xs[n & 1],
                        hs[(n - 1) & 1], hs[n & 1], c->dpart, (size_t)c->cap, n, bcg);
  HIPCHK(hipEventRecord(e1, c->sm));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the result of sweep nsweep is in buffer (nsweep & 1)
  if (nsweep & 1) {
    launch_copy<T>(c->sm, c->dactive, c->nactive, xs[0], hs[0], xs[1], hs[1]);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->sm));
  return 0;
}

int report(ffvhc_ctx* c, int* count, double* residual) {
  const std::vector<int> ids = selected_cases(c);
  HIPCHK(hipStreamSynchronize(c->sm));
  HIPCHK(hipMemcpy(c->hst.data(), c->dst, sizeof(SolveState) * c->cases.size(), hipMemcpyDeviceToHost));
  int status = 0;
  for (size_t i = 0; i < ids.size(); i++) {
    const SolveState& s = c->hst[ids[i]];
    if (count) count[i] = s.count;
    if (residual) residual[i] = s.residual;
    if (s.status) status = 1;
  }
  if (status) set_error("residual is NaN or Inf");
  return status;
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
  if (!c) return fail(-1, "null context");
  int prec = field_ok(c, x, 0) ? c->fields[x].prec : 8;
  int rc = check_args(c, x, A, b, prec);
  if (rc) return rc;
  if (bcg < 0 || bcg >= NUM_BC) return fail(-1, "bad bc group");
  rc = refresh_active(c);
  if (rc) return rc;
  if (countMax < 1) {
    // for(count=1; count<=countMax; ...) never runs: count = 1, residual untouched
    const std::vector<int> ids = selected_cases(c);
    for (size_t i = 0; i < ids.size(); i++) {
      if (count) count[i] = 1;
    }
    return 0;
  }
  rc = prec == 8 ? rbgs_run<double>(c, x, A, b, bcg, omega, countMax, eps)
                 : rbgs_run<float>(c, x, A, b, bcg, omega, countMax, eps);
  if (rc) return rc;
  return report(c, count, residual);
}

int ffvhc_rbgs_fixed(ffvhc_ctx* c, int x, const int* A, int b, int bcg, double omega, int nsweep, double* ms) {
  if (!c) return fail(-1, "null context");
  int prec = field_ok(c, x, 0) ? c->fields[x].prec : 8;
  int rc = check_args(c, x, A, b, prec);
  if (rc) return rc;
  if (bcg < 0 || bcg >= NUM_BC || nsweep < 1) return fail(-1, "bad bc group / sweep count");
  rc = refresh_active(c);
  if (rc) return rc;
  return prec == 8 ? rbgs_fixed_run<double>(c, x, A, b, bcg, omega, nsweep, ms)
                   : rbgs_fixed_run<float>(c, x, A, b, bcg, omega, nsweep, ms);
}

}  // extern "C"
