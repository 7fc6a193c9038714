// C ABI implementation: octree, context, cases, fields, ghost fill.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>

#include <rocblas/rocblas.h>

#include "kernels.h"
#include "runtime.h"

using namespace ffvhc;

namespace ffvhc {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

bool field_ok(ffvhc_ctx* ctx, int f, int prec) {
  // 16-byte records are solver scratch, never a caller's field
  return f >= 0 && f < (int)ctx->fields.size() && ctx->fields[f].used &&
         (prec == 0 ? ctx->fields[f].prec != 16 : ctx->fields[f].prec == prec);
}

std::vector<int> selected_cases(ffvhc_ctx* ctx) {
  std::vector<int> ids;
  if (!ctx->sel.empty()) {
    for (int c : ctx->sel)
      if (c >= 0 && c < (int)ctx->cases.size() && ctx->cases[c].used) ids.push_back(c);
    return ids;
  }
  for (int c = 0; c < (int)ctx->cases.size(); c++)
    if (ctx->cases[c].used) ids.push_back(c);
  return ids;
}

int refresh_active(ffvhc_ctx* ctx) {
  if (!ctx->dirty) return 0;
  // the launch-batch hints are iteration counts of the previous mix of cases: a new active set
  // starts without them (the batches change launches and host waits only, never results)
  ctx->bicg_hints.clear();
  ctx->hactive.clear();
  ctx->hcaseact.clear();
  ctx->multilevel = false;
  // the selected cases in selection order (as selected_cases lists them), without building the
  // list: the active lists have their full capacity reserved, so this allocates nothing
  auto add = [ctx](int c) {
    const CaseRec& r = ctx->cases[c];
    for (int s = 0; s < r.nblk; s++) ctx->hactive.push_back(r.slot_begin + s);
    if (r.nblk > 0) ctx->hcaseact.push_back(c);
    if (r.max_level > 0) ctx->multilevel = true;
  };
  if (!ctx->sel.empty()) {
    for (int c : ctx->sel)
      if (c >= 0 && c < (int)ctx->cases.size() && ctx->cases[c].used) add(c);
  } else {
    for (int c = 0; c < (int)ctx->cases.size(); c++)
      if (ctx->cases[c].used) add(c);
  }
  ctx->nactive = (int)ctx->hactive.size();
  ctx->ncaseact = (int)ctx->hcaseact.size();
  if (ctx->nactive)
    HIPCHK(hipMemcpyAsync(ctx->dactive, ctx->hactive.data(), sizeof(int) * ctx->nactive, hipMemcpyHostToDevice,
                          ctx->sm));
  if (ctx->ncaseact)
    HIPCHK(hipMemcpyAsync(ctx->dcaseact, ctx->hcaseact.data(), sizeof(int) * ctx->ncaseact, hipMemcpyHostToDevice,
                          ctx->sm));
  HIPCHK(hipStreamSynchronize(ctx->sm));
  ctx->dirty = false;
  return 0;
}

int alloc_field(ffvhc_ctx* ctx, int prec, int* field) {
  // 8 / 4: FP64 / FP32 fields; 16: one 16-byte record per cell (solver scratch only)
  if (prec != 8 && prec != 4 && prec != 16) return fail(-1, "precision must be 8 or 4");
  int id = -1;
  for (int i = 0; i < (int)ctx->fields.size(); i++)
    if (!ctx->fields[i].used) {
      id = i;
      break;
    }
  if (id < 0) {
    if ((int)ctx->fields.size() >= MAX_FIELDS) return fail(-3, "too many fields");
    id = (int)ctx->fields.size();
    ctx->fields.emplace_back();  // capacity reserved at creation: existing records never move
  }
  FieldRec& f = ctx->fields[id];
  f = FieldRec();
  f.prec = prec;
  size_t nd = (size_t)ctx->cap * NC * prec, nh = (size_t)ctx->cap * NH * prec;
  HIPCHK(hipMalloc(&f.d[0], nd));
  HIPCHK(hipMalloc(&f.h[0], nh));
  HIPCHK(hipMemsetAsync(f.d[0], 0, nd, ctx->sm));
  HIPCHK(hipMemsetAsync(f.h[0], 0, nh, ctx->sm));
  ctx->bytes += (long long)(nd + nh);
  f.used = true;
  *field = id;
  return 0;
}

int scratch_field(ffvhc_ctx* ctx, int prec, int role, int* field) {
  const int key = prec * 100 + role;
  for (auto& kv : ctx->scratch_ids)
    if (kv.first == key) {
      *field = kv.second;
      return 0;
    }
  int rc = alloc_field(ctx, prec, field);
  if (rc) return rc;
  ctx->scratch_ids.push_back({key, *field});
  return 0;
}

int ensure_second_buffer(ffvhc_ctx* ctx, int field) {
  FieldRec& f = ctx->fields[field];
  if (f.d[1]) return 0;
  size_t nd = (size_t)ctx->cap * NC * f.prec, nh = (size_t)ctx->cap * NH * f.prec;
  HIPCHK(hipMalloc(&f.d[1], nd));
  HIPCHK(hipMalloc(&f.h[1], nh));
  HIPCHK(hipMemsetAsync(f.d[1], 0, nd, ctx->sm));
  HIPCHK(hipMemsetAsync(f.h[1], 0, nh, ctx->sm));
  ctx->bytes += (long long)(nd + nh);
  return 0;
}

int upload_case_meta(ffvhc_ctx* c, int id) {
  const CaseRec& r = c->cases[id];
  DevCase& d = c->hcase[id];
  d.slot_begin = r.slot_begin;
  d.nblk = r.nblk;
  d.rho = r.params.rho;
  d.mu = r.params.mu;
  d.dt = r.params.dt;
  d.gx = r.params.gx;
  d.gy = r.params.gy;
  d.gz = r.params.gz;
  d.betag = r.params.betag;
  d.tr = r.params.tr;
  for (int g = 0; g < NUM_USER_BC; g++)
    for (int f = 0; f < 6; f++) {
      d.bct[g][f] = r.params.bct[g][f];
      d.bcv[g][f] = r.params.bcv[g][f];
    }
  for (int f = 0; f < 6; f++) {
    d.bct[BC_HOM][f] = r.params.bct[BC_P][f];
    d.bcv[BC_HOM][f] = 0.0;
  }
  d.step = r.step;
  // Solver.cpp:3347-3355: the X_M inflow value of UX ramps up over the acceleration steps
  if (r.accel_steps > 0 && r.step <= r.accel_steps)
    d.bcv[BC_UX][0] = r.params.bcv[BC_UX][0] * (double)r.step / (double)r.accel_steps;
  HIPCHK(hipMemcpyAsync(c->dcase + id, &d, sizeof(DevCase), hipMemcpyHostToDevice, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  return 0;
}

}  // namespace ffvhc

extern "C" {

const char* ffvhc_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------------------------- tree

int ffvhc_tree_build(const ffvhc_tree_spec* spec, ffvhc_tree** out) {
  if (!spec || !out) return fail(-1, "null argument");
  if (spec->nregions < 0 || spec->nregions > FFVHC_MAX_REGIONS) return fail(-1, "bad region count");
  OctreeSpec s;
  for (int d = 0; d < 3; d++) {
    s.root[d] = spec->root[d];
    s.origin[d] = spec->origin[d];
    s.periodic[d] = spec->periodic[d];
  }
  s.root_length = spec->root_length;
  s.n = spec->n;
  s.min_level = spec->min_level;
  s.margin = spec->margin;
  s.ordering = spec->ordering;
  for (int r = 0; r < spec->nregions; r++) {
    const ffvhc_region& g = spec->regions[r];
    Region q{};
    if (g.kind < 0 || g.kind > 2) return fail(-1, "bad region kind");
    q.kind = (Region::Kind)g.kind;
    for (int d = 0; d < 3; d++) {
      q.a[d] = g.a[d];
      q.b[d] = g.b[d];
    }
    q.r = g.r;
    q.dr = g.dr;
    q.level = g.level;
    s.regions.push_back(q);
  }
  try {
    auto* t = new ffvhc_tree;
    t->tree = new Octree(s);
    *out = t;
  } catch (const std::exception& e) {
    return fail(-1, e.what());
  }
  return 0;
}

int ffvhc_tree_num_blocks(const ffvhc_tree* t) { return t ? (int)t->tree->blocks().size() : -1; }
int ffvhc_tree_max_level(const ffvhc_tree* t) { return t ? t->tree->max_level() : -1; }

int ffvhc_tree_table(const ffvhc_tree* t, int* level, double* org, double* dx, int* ftype, int* nbr, int* sub) {
  if (!t) return fail(-1, "null tree");
  const auto& B = t->tree->blocks();
  for (size_t i = 0; i < B.size(); i++) {
    if (level) level[i] = B[i].level;
    if (dx) dx[i] = B[i].dx;
    for (int d = 0; d < 3; d++)
      if (org) org[3 * i + d] = B[i].org[d];
    for (int f = 0; f < 6; f++) {
      if (ftype) ftype[6 * i + f] = B[i].ftype[f];
      if (sub) sub[6 * i + f] = B[i].sub[f];
      for (int s = 0; s < 4; s++)
        if (nbr) nbr[24 * i + 4 * f + s] = B[i].nbr[f][s];
    }
  }
  return 0;
}

void ffvhc_tree_free(ffvhc_tree* t) {
  if (!t) return;
  delete t->tree;
  delete t;
}

// ------------------------------------------------------------------ allocation-free scratch
// Device memory of the allocation-free entry point (ffvhc_force_p_noalloc): the per-block partial
// triples of up to FFVHC_NOALLOC_SLOTS slots, then the per-case triples of up to MAX_CASES cases
// (3.2 MB).  Every device the library runs on holds its own copy until the process ends.  In host
// code `scratch` is a host-side stand-in, never a device address: the address on a device comes
// from hipGetSymbolAddress with that device current, looked up once per device by the first
// context created on it and kept in g_scratch_dev.
constexpr size_t SCRATCH_DOUBLES = 3 * (size_t)FFVHC_NOALLOC_SLOTS + 3 * (size_t)MAX_CASES;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static double* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// the current device's scratch address (the caller has made `device` current)
static int scratch_address(int device, double** out) {
  if (device >= SCRATCH_MAX_DEVICES) return fail(-1, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(-2, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_dev[device] = (double*)p;
  }
  *out = g_scratch_dev[device];
  return 0;
}

// ---------------------------------------------------------------------------------- context

int ffvhc_ctx_create(int device, long long max_blocks, ffvhc_ctx** out) {
  if (!out || max_blocks <= 0) return fail(-1, "bad arguments");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(-1, "no such device");
  HIPCHK(hipSetDevice(device));
  std::unique_ptr<ffvhc_ctx> c(new ffvhc_ctx);
  c->device = device;
  c->cap = max_blocks;
  HIPCHK(hipStreamCreateWithFlags(&c->sm, hipStreamNonBlocking));
  c->hblk.assign((size_t)max_blocks, DevBlock{});
  c->hcase.assign(MAX_CASES, DevCase{});
  c->hst.assign(MAX_CASES, SolveState{});
  c->hst2.assign(MAX_CASES, SolveState{});
  c->fields.reserve(MAX_FIELDS);
  HIPCHK(hipMalloc(&c->dblk, sizeof(DevBlock) * max_blocks));
  HIPCHK(hipMalloc(&c->dcase, sizeof(DevCase) * MAX_CASES));
  HIPCHK(hipMalloc(&c->dactive, sizeof(int) * max_blocks));
  HIPCHK(hipMalloc(&c->dcaseact, sizeof(int) * MAX_CASES));
  HIPCHK(hipMalloc(&c->dst, sizeof(SolveState) * MAX_CASES));
  HIPCHK(hipMalloc(&c->dst2, sizeof(SolveState) * MAX_CASES));
  HIPCHK(hipMalloc(&c->dpart, sizeof(double) * 6 * max_blocks));
  HIPCHK(hipMalloc(&c->dcut, sizeof(float) * 6 * NC * max_blocks));
  HIPCHK(hipMalloc(&c->dflag, sizeof(uint32_t) * NC * max_blocks));
  HIPCHK(hipMalloc(&c->dcid, sizeof(uint8_t) * 6 * NC * max_blocks));
  HIPCHK(hipMemset(c->dst, 0, sizeof(SolveState) * MAX_CASES));
  HIPCHK(hipMemset(c->dst2, 0, sizeof(SolveState) * MAX_CASES));
  int rc = scratch_address(device, &c->dscratch);
  if (rc) return rc;
  // the active lists never grow past these, so refreshing them allocates no host memory
  c->hactive.reserve((size_t)max_blocks);
  c->hcaseact.reserve(MAX_CASES);
  c->bytes = (long long)(sizeof(DevBlock) + sizeof(int) + 4 * sizeof(double) + (6 * 4 + 4 + 6) * NC) * max_blocks;
  c->free_slots.push_back({0, max_blocks});
  *out = c.release();
  return 0;
}

void ffvhc_ctx_destroy(ffvhc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->sm);
  if (c->sc) (void)hipStreamSynchronize(c->sc);
  for (auto& f : c->fields)
    for (int b = 0; b < 2; b++) {
      if (f.d[b]) (void)hipFree(f.d[b]);
      if (f.h[b]) (void)hipFree(f.h[b]);
    }
  (void)hipFree(c->dblk);
  (void)hipFree(c->dcase);
  (void)hipFree(c->dactive);
  (void)hipFree(c->dcaseact);
  (void)hipFree(c->dst);
  (void)hipFree(c->dst2);
  (void)hipFree(c->dpart);
  (void)hipFree(c->dcm);
  (void)hipFree(c->dci);
  (void)hipFree(c->dcoff);
  (void)hipFree(c->dcsnap);
  (void)hipFree(c->dce);
  (void)hipFree(c->dcuni);
  (void)hipFree(c->dctab);
  (void)hipFree(c->dctrec);
  (void)hipFree(c->dcut);
  (void)hipFree(c->dflag);
  (void)hipFree(c->dcid);
  if (c->blas) (void)rocblas_destroy_handle((rocblas_handle)c->blas);
  (void)hipStreamDestroy(c->sm);
  if (c->sc) {
    (void)hipEventDestroy(c->ev_cs);
    (void)hipEventDestroy(c->ev_cd);
    (void)hipStreamDestroy(c->sc);
  }
  delete c;
}

int ffvhc_ctx_sync(ffvhc_ctx* c) {
  if (!c) return fail(-1, "null context");
  HIPCHK(hipStreamSynchronize(c->sm));
  return 0;
}

long long ffvhc_ctx_device_bytes(const ffvhc_ctx* c) { return c ? c->bytes : -1; }

int ffvhc_ctx_select(ffvhc_ctx* c, const int* ids, int n) {
  if (!c) return fail(-1, "null context");
  for (int i = 0; i < n; i++)
    if (ids[i] < 0 || ids[i] >= (int)c->cases.size() || !c->cases[ids[i]].used) return fail(-4, "no such case");
  // the same selection again keeps the active lists (and with them the launch-batch hints)
  if (!c->dirty && (int)c->sel.size() == n && std::equal(c->sel.begin(), c->sel.end(), ids)) return 0;
  c->sel.assign(ids, ids + n);
  c->dirty = true;
  return refresh_active(c);
}

// ------------------------------------------------------------------------------------ cases


int ffvhc_case_add(ffvhc_ctx* c, const ffvhc_tree* t, const ffvhc_case_params* p, int g, const double* cut,
                   const int* cid, const int* pid, int* case_id) {
  if (!c || !t || !p || !case_id) return fail(-1, "null argument");
  if (t->tree->spec().n != NB) return fail(-1, "device blocks are 16^3 cells");
  if ((cut || cid || pid) && g < 1) return fail(-1, "geometry needs g >= 1");
  c->geom_serial++;
  const auto& B = t->tree->blocks();
  const int64_t nb = (int64_t)B.size();
  // first-fit contiguous slot range
  int64_t begin = -1;
  for (auto& iv : c->free_slots)
    if (iv.second - iv.first >= nb) {
      begin = iv.first;
      iv.first += nb;
      break;
    }
  if (begin < 0) return fail(-3, "not enough free block slots");
  c->free_slots.erase(std::remove_if(c->free_slots.begin(), c->free_slots.end(),
                                     [](const std::pair<int64_t, int64_t>& iv) { return iv.first >= iv.second; }),
                      c->free_slots.end());
  int id = -1;
  for (int i = 0; i < (int)c->cases.size(); i++)
    if (!c->cases[i].used) {
      id = i;
      break;
    }
  if (id < 0) {
    if ((int)c->cases.size() >= MAX_CASES) return fail(-3, "too many cases");
    id = (int)c->cases.size();
    c->cases.emplace_back();
  }
  CaseRec& r = c->cases[id];
  r.used = true;
  r.slot_begin = (int)begin;
  r.nblk = (int)nb;
  r.max_level = t->tree->max_level();
  r.root_length = t->tree->spec().root_length;
  r.ncells = nb * NC;
  r.params = *p;
  // block table with global slot numbers
  std::vector<float> hcut;
  std::vector<uint32_t> hflag;
  std::vector<uint8_t> hcidv;
  const int np = NB + 2 * g;
  const size_t npad = (size_t)np * np * np;
  hcut.assign((size_t)6 * NC * nb, 1.0f);
  hflag.assign((size_t)NC * nb, 1u << 6);
  hcidv.assign((size_t)6 * NC * nb, 0);
  for (int64_t i = 0; i < nb; i++) {
    DevBlock& d = c->hblk[begin + i];
    d = DevBlock{};
    d.case_id = id;
    d.level = B[i].level;
    d.dx = B[i].dx;
    for (int q = 0; q < 3; q++) d.org[q] = B[i].org[q];
    for (int f = 0; f < 6; f++) {
      d.ftype[f] = B[i].ftype[f];
      d.sub[f] = B[i].sub[f];
      for (int s = 0; s < 4; s++) d.nbr[f][s] = B[i].nbr[f][s] < 0 ? -1 : (int)(begin + B[i].nbr[f][s]);
    }
    int irregular = 0;
    if (cut || cid || pid) {
      for (int k = 0; k < NB; k++)
        for (int j = 0; j < NB; j++)
          for (int ii = 0; ii < NB; ii++) {
            size_t m = (size_t)i * npad + (size_t)(ii + g) + np * ((size_t)(j + g) + np * (size_t)(k + g));
            size_t q = (size_t)i * NC + ii + NB * (j + NB * k);
            uint32_t fl = 0;
            for (int f = 0; f < 6; f++) {
              int id6 = cid ? cid[(size_t)f * nb * npad + m] : 0;
              if (id6 != 0) {
                fl |= 1u << f;
                if (id6 < 0 || id6 > 255) return fail(-1, "cut ids must be in 0..255");
              }
              hcidv[(size_t)f * NC * nb + q] = (uint8_t)id6;
              float cv = cut ? (float)cut[(size_t)f * nb * npad + m] : 1.0f;
              if (cut && (double)cv != cut[(size_t)f * nb * npad + m])
                return fail(-1, "cut distances must be float32 values (Cutlib CutPos32Array)");
              hcut[(size_t)f * NC * nb + q] = cv;
            }
            if (!pid || pid[m] == 1)
              fl |= 1u << 6;
            hflag[q] = fl;
            if (fl != (1u << 6)) irregular = 1;
          }
    }
    d.flags = irregular;
  }
  // geometry is stored per global slot: [f][slot][cell] for cuts/cids, [slot][cell] for flags
  HIPCHK(hipMemcpyAsync(c->dblk + begin, c->hblk.data() + begin, sizeof(DevBlock) * nb, hipMemcpyHostToDevice, c->sm));
  for (int f = 0; f < 6; f++) {
    HIPCHK(hipMemcpyAsync(c->dcut + (size_t)f * NC * c->cap + (size_t)begin * NC, hcut.data() + (size_t)f * NC * nb,
                          sizeof(float) * NC * nb, hipMemcpyHostToDevice, c->sm));
    HIPCHK(hipMemcpyAsync(c->dcid + (size_t)f * NC * c->cap + (size_t)begin * NC, hcidv.data() + (size_t)f * NC * nb,
                          NC * nb, hipMemcpyHostToDevice, c->sm));
  }
  HIPCHK(hipMemcpyAsync(c->dflag + (size_t)begin * NC, hflag.data(), sizeof(uint32_t) * NC * nb,
                        hipMemcpyHostToDevice, c->sm));
  int rc = upload_case_meta(c, id);
  if (rc) return rc;
  c->dirty = true;
  *case_id = id;
  return refresh_active(c);
}

int ffvhc_case_fill_phase(ffvhc_ctx* c, int id, const double* seed, int cut_id, long long* counts) {
  if (!c || !seed) return fail(-1, "null argument");
  if (id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  if (cut_id < 1 || cut_id > 255) return fail(-1, "cut_id must be in 1..255");
  const CaseRec& r = c->cases[id];
  const int nb = r.nblk, b0 = r.slot_begin;
  // the seed cell: the first block (in block order) whose cells hold the point, as the host
  // geometry's search starts
  int sb = -1, si = 0, sj = 0, sk = 0;
  for (int b = 0; b < nb && sb < 0; b++) {
    const DevBlock& B = c->hblk[b0 + b];
    int q[3];
    bool in = true;
    for (int a = 0; a < 3; a++) {
      q[a] = (int)std::floor((seed[a] - B.org[a]) / B.dx);
      if (q[a] < 0 || q[a] >= NB) in = false;
    }
    if (in) {
      sb = b;
      si = q[0];
      sj = q[1];
      sk = q[2];
    }
  }
  if (sb < 0) return fail(-1, "the seed point lies outside the case's blocks");
  std::vector<int> slots(nb);
  for (int b = 0; b < nb; b++) slots[b] = b0 + b;
  std::vector<uint16_t> hlab((size_t)nb * NF, 0);
  hlab[(size_t)sb * NF + si + NB * sj] = (uint16_t)(1u << sk);
  constexpr int BATCH = 8;
  int* dslots = nullptr;
  int* dflags = nullptr;  // changed[BATCH], then nfluid[cap], irregular[cap]
  uint16_t* dlab = nullptr;
  HIPCHK(hipMallocAsync((void**)&dslots, sizeof(int) * nb, c->sm));
  HIPCHK(hipMallocAsync((void**)&dflags, sizeof(int) * (BATCH + 2 * (size_t)c->cap), c->sm));
  HIPCHK(hipMallocAsync((void**)&dlab, sizeof(uint16_t) * NF * (size_t)c->cap, c->sm));
  HIPCHK(hipMemcpyAsync(dslots, slots.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c->sm));
  HIPCHK(hipMemcpyAsync(dlab + (size_t)b0 * NF, hlab.data(), sizeof(uint16_t) * NF * nb, hipMemcpyHostToDevice,
                        c->sm));
  const size_t capnc = (size_t)c->cap * NC;
  int hchg[BATCH];
  bool done = false;
  for (int pass = 0; !done; pass += BATCH) {
    if (pass > 64 * nb * NB) return fail(-2, "phase flood fill did not settle");
    HIPCHK(hipMemsetAsync(dflags, 0, sizeof(int) * BATCH, c->sm));
    for (int q = 0; q < BATCH; q++)
      launch_phase_pass(c->sm, dslots, nb, c->dblk, c->dcid, capnc, cut_id, dlab, dflags + q);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hchg, dflags, sizeof(hchg), hipMemcpyDeviceToHost, c->sm));
    HIPCHK(hipStreamSynchronize(c->sm));
    done = hchg[BATCH - 1] == 0;  // a pass that changed nothing: every later one is the same
  }
  int* dnf = dflags + BATCH;
  int* dirr = dnf + c->cap;
  launch_phase_store(c->sm, dslots, nb, dlab, c->dflag, dnf, dirr);
  HIPCHK(hipGetLastError());
  std::vector<int> nf(nb), irr(nb);
  HIPCHK(hipMemcpyAsync(nf.data(), dnf + b0, sizeof(int) * nb, hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipMemcpyAsync(irr.data(), dirr + b0, sizeof(int) * nb, hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipFreeAsync(dslots, c->sm));
  HIPCHK(hipFreeAsync(dflags, c->sm));
  HIPCHK(hipFreeAsync(dlab, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  long long fluid = 0;
  for (int b = 0; b < nb; b++) {
    fluid += nf[b];
    c->hblk[b0 + b].flags = irr[b];
  }
  HIPCHK(hipMemcpyAsync(c->dblk + b0, c->hblk.data() + b0, sizeof(DevBlock) * nb, hipMemcpyHostToDevice, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  c->geom_serial++;  // the pressure matrix depends on the phase
  if (counts) {
    counts[0] = fluid;
    counts[1] = (long long)nb * NC - fluid;
  }
  return 0;
}

int ffvhc_case_download_phase(ffvhc_ctx* c, int id, int g, int* pid) {
  if (!c || !pid || g < 0) return fail(-1, "bad arguments");
  if (id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  const CaseRec& r = c->cases[id];
  std::vector<uint32_t> fl((size_t)r.nblk * NC);
  HIPCHK(hipMemcpyAsync(fl.data(), c->dflag + (size_t)r.slot_begin * NC, sizeof(uint32_t) * fl.size(),
                        hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  const int np = NB + 2 * g;
  const size_t npad = (size_t)np * np * np;
  std::fill(pid, pid + (size_t)r.nblk * npad, 0);
  for (int b = 0; b < r.nblk; b++)
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++)
          pid[(size_t)b * npad + (size_t)(i + g) + (size_t)np * ((size_t)(j + g) + (size_t)np * (size_t)(k + g))] =
              (int)((fl[(size_t)b * NC + i + NB * (j + NB * k)] >> 6) & 1u);
  return 0;
}

int ffvhc_case_remove(ffvhc_ctx* c, int id) {
  if (!c || id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  CaseRec& r = c->cases[id];
  c->free_slots.push_back({r.slot_begin, (int64_t)r.slot_begin + r.nblk});
  std::sort(c->free_slots.begin(), c->free_slots.end());
  std::vector<std::pair<int64_t, int64_t>> merged;
  for (auto& iv : c->free_slots) {
    if (!merged.empty() && merged.back().second == iv.first) merged.back().second = iv.second;
    else merged.push_back(iv);
  }
  c->free_slots = merged;
  r = CaseRec();
  c->sel.erase(std::remove(c->sel.begin(), c->sel.end(), id), c->sel.end());
  c->dirty = true;
  c->geom_serial++;
  return refresh_active(c);
}

int ffvhc_case_num_blocks(const ffvhc_ctx* c, int id) {
  if (!c || id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return -1;
  return c->cases[id].nblk;
}

int ffvhc_case_set_bc(ffvhc_ctx* c, int id, int group, int face, int type, double value) {
  if (!c || id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  if (group < 0 || group >= NUM_USER_BC || face < 0 || face > 5) return fail(-1, "bad group/face");
  c->cases[id].params.bct[group][face] = type;
  c->cases[id].params.bcv[group][face] = value;
  c->geom_serial++;
  return upload_case_meta(c, id);
}

// ----------------------------------------------------------------------------------- fields

int ffvhc_field_alloc(ffvhc_ctx* c, int prec, int* field) {
  if (!c || !field) return fail(-1, "null argument");
  if (prec != 8 && prec != 4) return fail(-1, "precision must be 8 or 4");
  return alloc_field(c, prec, field);
}

int ffvhc_field_free(ffvhc_ctx* c, int f) {
  if (!c || !field_ok(c, f, 0)) return fail(-4, "no such field");
  HIPCHK(hipStreamSynchronize(c->sm));
  FieldRec& r = c->fields[f];
  for (int b = 0; b < 2; b++) {
    if (r.d[b]) (void)hipFree(r.d[b]);
    if (r.h[b]) (void)hipFree(r.h[b]);
  }
  r = FieldRec();
  return 0;
}

}  // extern "C"

template <class T>
static void pack(const T* pad, int g, T* in, T* halo, int64_t nb) {
  const int np = NB + 2 * g;
  const size_t npad = (size_t)np * np * np;
  for (int64_t b = 0; b < nb; b++) {
    const T* P = pad + b * npad;
    auto at = [&](int i, int j, int k) { return P[(size_t)(i + g) + np * ((size_t)(j + g) + np * (size_t)(k + g))]; };
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++) in[b * NC + i + NB * (j + NB * k)] = at(i, j, k);
    for (int f = 0; f < 6; f++)
      for (int bb = 0; bb < NB; bb++)
        for (int a = 0; a < NB; a++) {
          int i, j, k;
          face_cell(f, a, bb, -1, i, j, k);
          halo[b * NH + halo_index(f, a, bb)] = at(i, j, k);
        }
  }
}

template <class T>
static void unpack(const T* in, const T* halo, int g, T* pad, int64_t nb) {
  const int np = NB + 2 * g;
  const size_t npad = (size_t)np * np * np;
  for (int64_t b = 0; b < nb; b++) {
    T* P = pad + b * npad;
    auto at = [&](int i, int j, int k) -> T& {
      return P[(size_t)(i + g) + np * ((size_t)(j + g) + np * (size_t)(k + g))];
    };
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++) at(i, j, k) = in[b * NC + i + NB * (j + NB * k)];
    for (int f = 0; f < 6; f++)
      for (int bb = 0; bb < NB; bb++)
        for (int a = 0; a < NB; a++) {
          int i, j, k;
          face_cell(f, a, bb, -1, i, j, k);
          at(i, j, k) = halo[b * NH + halo_index(f, a, bb)];
        }
  }
}

extern "C" {

int ffvhc_field_upload(ffvhc_ctx* c, int f, int id, int g, const void* padded) {
  if (!c || !field_ok(c, f, 0) || !padded || g < 1) return fail(-1, "bad arguments");
  if (id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  const CaseRec& r = c->cases[id];
  FieldRec& F = c->fields[f];
  const int64_t nb = r.nblk;
  std::vector<char> in((size_t)nb * NC * F.prec), halo((size_t)nb * NH * F.prec);
  if (F.prec == 8) pack((const double*)padded, g, (double*)in.data(), (double*)halo.data(), nb);
  else pack((const float*)padded, g, (float*)in.data(), (float*)halo.data(), nb);
  HIPCHK(hipMemcpyAsync((char*)F.d[0] + (size_t)r.slot_begin * NC * F.prec, in.data(), in.size(),
                        hipMemcpyHostToDevice, c->sm));
  HIPCHK(hipMemcpyAsync((char*)F.h[0] + (size_t)r.slot_begin * NH * F.prec, halo.data(), halo.size(),
                        hipMemcpyHostToDevice, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  return 0;
}

int ffvhc_field_download(ffvhc_ctx* c, int f, int id, int g, void* padded) {
  if (!c || !field_ok(c, f, 0) || !padded || g < 1) return fail(-1, "bad arguments");
  if (id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  const CaseRec& r = c->cases[id];
  FieldRec& F = c->fields[f];
  const int64_t nb = r.nblk;
  std::vector<char> in((size_t)nb * NC * F.prec), halo((size_t)nb * NH * F.prec);
  HIPCHK(hipMemcpyAsync(in.data(), (char*)F.d[0] + (size_t)r.slot_begin * NC * F.prec, in.size(),
                        hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipMemcpyAsync(halo.data(), (char*)F.h[0] + (size_t)r.slot_begin * NH * F.prec, halo.size(),
                        hipMemcpyDeviceToHost, c->sm));
  HIPCHK(hipStreamSynchronize(c->sm));
  if (F.prec == 8) unpack((const double*)in.data(), (const double*)halo.data(), g, (double*)padded, nb);
  else unpack((const float*)in.data(), (const float*)halo.data(), g, (float*)padded, nb);
  return 0;
}

/* LocalScalar3D<real>::Dump2 (LocalScalar3D.cpp:215-264) file layout: six ints {nx, ny, nz, vc,
 * nE = 1, nB} then, per block in block order, the padded (n+2vc)^3 FP64 array. */
int ffvhc_field_dump(ffvhc_ctx* c, int f, int id, int g, int bcg, int updater, const char* path) {
  if (!c || !path || g < 1 || g > 2) return fail(-1, "bad arguments (ghost width 1 or 2)");
  if (bcg < 0 || bcg >= NUM_BC || (updater != 1 && updater != 5)) return fail(-1, "bad bc group / updater");
  if (!field_ok(c, f, 8)) return fail(-1, "Dump2 writes FP64 fields");
  if (id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  const int nb = c->cases[id].nblk;
  const size_t npad = (size_t)(NB + 2 * g) * (NB + 2 * g) * (NB + 2 * g);
  std::vector<double> data((size_t)nb * npad, 0.0);
  int rc = ffvhc_field_download(c, f, id, g, data.data());
  if (rc) return rc;
  // Dump2 fills the ghosts first (ImposeBoundaryCondition): every ghost layer, rebuilt on the
  // host from the interiors
  rc = host_impose_bc(c, id, data.data(), g, bcg, updater);
  if (rc) return rc;
  FILE* fp = fopen(path, "wb");
  if (!fp) return fail(-1, std::string("cannot open ") + path);
  const int head[6] = {NB, NB, NB, g, 1, nb};
  bool ok = fwrite(head, sizeof(int), 6, fp) == 6 && fwrite(data.data(), sizeof(double), data.size(), fp) == data.size();
  ok = (fclose(fp) == 0) && ok;
  return ok ? 0 : fail(-1, std::string("write failed: ") + path);
}

/* Header of a Dump2 file: {nx, ny, nz, vc, nE, nB}. */
int ffvhc_dump_header(const char* path, int* head6) {
  if (!path || !head6) return fail(-1, "bad arguments");
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(-1, std::string("cannot open ") + path);
  const bool ok = fread(head6, sizeof(int), 6, fp) == 6;
  fclose(fp);
  return ok ? 0 : fail(-1, std::string("short header: ") + path);
}

/* LocalScalar3D<real>::Load2 (LocalScalar3D.cpp:266-342): the file must hold nE = 1 and the
 * case's block count.  Same resolution (16^3 blocks): the interior and the first ghost layer are
 * uploaded.  Half resolution (8^3 blocks, LocalScalar3D.cpp:313-331): each fine cell takes the
 * value of the coarse cell that holds it, x(i, j, k) = xc((i+1)/2, (j+1)/2, (k+1)/2) in 1-based
 * indices (sf3d_copy_x2, bsf3d.f90:5-33); the ghosts are left for the caller's fill, as Load2
 * leaves them to ImposeBoundaryCondition. */
int ffvhc_field_load(ffvhc_ctx* c, int f, int id, const char* path) {
  if (!c || !path) return fail(-1, "bad arguments");
  if (!field_ok(c, f, 8)) return fail(-1, "Load2 reads FP64 fields");
  if (id < 0 || id >= (int)c->cases.size() || !c->cases[id].used) return fail(-4, "no such case");
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(-1, std::string("cannot open ") + path);
  int head[6];
  if (fread(head, sizeof(int), 6, fp) != 6) {
    fclose(fp);
    return fail(-1, "short header");
  }
  const int nb = c->cases[id].nblk, g = head[3];
  const bool same = head[0] == NB && head[1] == NB && head[2] == NB;
  const bool half = head[0] == NB / 2 && head[1] == NB / 2 && head[2] == NB / 2;
  if (!(same || half) || head[4] != 1 || head[5] != nb || g < 1 || g > 8) {
    fclose(fp);
    return fail(-1, "file does not match the case (block size 16^3 or 8^3, nE, block count)");
  }
  const int nf = same ? NB : NB / 2, npf = nf + 2 * g, np = NB + 2 * g;
  const size_t nfile = (size_t)npf * npf * npf, npad = (size_t)np * np * np;
  std::vector<double> data((size_t)nb * nfile);
  const bool ok = fread(data.data(), sizeof(double), data.size(), fp) == data.size();
  fclose(fp);
  if (!ok) return fail(-1, "short file");
  if (same) return ffvhc_field_upload(c, f, id, g, data.data());
  std::vector<double> fine((size_t)nb * npad, 0.0);
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < NB; k++)
      for (int j = 0; j < NB; j++)
        for (int i = 0; i < NB; i++)
          fine[(size_t)b * npad + (size_t)(i + g) + (size_t)np * ((size_t)(j + g) + (size_t)np * (size_t)(k + g))] =
              data[(size_t)b * nfile + (size_t)(i / 2 + g) +
                   (size_t)npf * ((size_t)(j / 2 + g) + (size_t)npf * (size_t)(k / 2 + g))];
  return ffvhc_field_upload(c, f, id, g, fine.data());
}

int ffvhc_ghost_fill(ffvhc_ctx* c, int f, int bcg, int updater) {
  if (!c || !field_ok(c, f, 0)) return fail(-1, "bad field");
  if (bcg < 0 || bcg >= NUM_BC || (updater != 1 && updater != 5)) return fail(-1, "bad bc group / updater");
  int rc = refresh_active(c);
  if (rc) return rc;
  FieldRec& F = c->fields[f];
  if (F.prec == 8)
    launch_ghost_pull<double>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const double*)F.d[0],
                              (double*)F.h[0], bcg, updater, c->multilevel);
  else
    launch_ghost_pull<float>(c->sm, c->dactive, c->nactive, c->dblk, c->dcase, (const float*)F.d[0],
                             (float*)F.h[0], bcg, updater, c->multilevel);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"
