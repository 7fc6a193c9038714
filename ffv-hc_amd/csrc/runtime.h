// Host-side runtime state behind the C ABI: one context per GPU holding every resident case's
// blocks in device slots, the fields allocated over all slots, and solver scratch.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/ffvhc.h"
#include "device.h"
#include "kernels.h"
#include "octree.h"

struct ffvhc_tree {
  ffvhc::Octree* tree;
};

namespace ffvhc {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return ffvhc::fail(-2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct FieldRec {
  bool used = false;
  int prec = 8;          // bytes per value
  void* d[2] = {nullptr, nullptr};  // interior, ping-pong
  void* h[2] = {nullptr, nullptr};  // halo, ping-pong
};

struct CaseRec {
  bool used = false;
  int slot_begin = 0;
  int nblk = 0;
  int max_level = 0;
  double root_length = 1.0;
  int64_t ncells = 0;
  int step = 0;         // Solver::Update step counter
  int accel_steps = 0;  // TimeControl/Acceleration of the X_M inflow (0 = off)
  ffvhc_case_params params{};
};

}  // namespace ffvhc

struct ffvhc_ctx {
  int device = 0;
  hipStream_t sm = nullptr;
  int64_t cap = 0;  // slots
  // block table
  std::vector<ffvhc::DevBlock> hblk;
  ffvhc::DevBlock* dblk = nullptr;
  std::vector<ffvhc::DevCase> hcase;
  ffvhc::DevCase* dcase = nullptr;
  std::vector<ffvhc::CaseRec> cases;
  // geometry per slot: cut distances (float, exact: Cutlib stores float32), cell flags
  // (bits 0-5 face d cut, bit 6 fluid), cut ids (uint8 per face, force kernel only)
  float* dcut = nullptr;
  uint32_t* dflag = nullptr;
  uint8_t* dcid = nullptr;
  // active block list (selected cases)
  std::vector<int> sel;       // selected case ids (empty = all)
  std::vector<int> hactive;
  int* dactive = nullptr;
  int nactive = 0;
  std::vector<int> hcaseact;  // selected case ids with at least one block
  int* dcaseact = nullptr;
  int ncaseact = 0;
  bool multilevel = false;
  bool dirty = true;
  // solver state
  ffvhc::SolveState* dst = nullptr;   // inner / single-level solve state per case
  ffvhc::SolveState* dst2 = nullptr;  // outer (defect correction) state per case
  std::vector<ffvhc::SolveState> hst;
  std::vector<ffvhc::SolveState> hst2;
  // internally allocated scratch fields, keyed by precision * 100 + role
  std::vector<std::pair<int, int>> scratch_ids;
  double* dpart = nullptr;  // 6 * cap per-block partials
  // this device's copy of the allocation-free scratch array (its address from
  // hipGetSymbolAddress at context creation)
  double* dscratch = nullptr;
  // coarse correction of the two-level preconditioners: dense A_c per case (inverted in place),
  // its FP32 copy (row pitch padded to a multiple of 4), per-case offsets of both (dcoff[0..MAX_CASES)
  // and dcoff[MAX_CASES..)), Gauss-Jordan panel scratch, corrections and block sums per slot
  double* dcm = nullptr;
  size_t dcm_n = 0;
  float* dci = nullptr;
  size_t dci_n = 0;
  long long* dcoff = nullptr;
  double* dcsnap = nullptr;
  size_t dcsnap_n = 0;
  double* dce = nullptr;
  // the coarse inverse of n >= coarse_gemm_min blocks runs its panel products as FP64 GEMMs
  // (rocBLAS handle `blas`, made on first use); smaller ones use the 32-wide VALU Gauss-Jordan
  void* blas = nullptr;
  int coarse_gemm_min = 2048;
  // host copies the setups read back or upload asynchronously; they live here so that they
  // outlive the copies (a setup returns without waiting for its last kernels)
  std::vector<double> hcrows;   // 25 coarse row entries per slot
  std::vector<long long> hcoff;  // the offsets uploaded to dcoff
  std::vector<int> hcuni;        // class-uniform flag per slot
  // the coarse inversion runs on stream sc (default priority) beside the solver's first kernels
  // on sm: ev_cs orders it after the assembly on sm, ev_cd marks its end; coarse_pending says
  // that sm has not yet been made to wait for ev_cd
  hipStream_t sc = nullptr;
  hipEvent_t ev_cs = nullptr, ev_cd = nullptr;
  bool coarse_pending = false;
  // setup_serial counts the writes of the solver setup buffers (FP32 band copies, preconditioner
  // records, class tables, coarse inverse); geom_serial counts the changes of what a time step's
  // pressure matrix is made from (cases added or removed, boundary conditions).  The time step
  // keeps the setup of its pressure solve (ffvhc_solve_mixed methods 11-28) while neither has
  // moved and the method, matrix fields, active blocks and options are the same
  // (FFVHC_OPTION_SETUP_REUSE).
  bool reuse_setup = true;
  // the multiplicative methods form an iteration's finalisation inside the next coarse application
  bool fuse_fin = true;
  bool step_solve = false;  // set by the time step around its pressure solve
  // the step's reference-pressure point (bcut_set_reference_value rewrites the rows of the cell
  // holding it and of its neighbours, so the matrix depends on it)
  bool step_ref_active = false;
  double step_ref[3] = {0.0, 0.0, 0.0};
  // first launch batch of a BiCGSTAB round: the iterations round r of the previous solve of the
  // same kind needed (key: the ffvhc_solve_mixed method, or -100 / -200 less the BC group for
  // ffvhc_bicgstab in FP64 / FP32); bicg_round is the round running, or -1 (no hint)
  int bicg_round = -1;
  int bicg_hint_key = 0;
  std::map<int, std::array<int, 4>> bicg_hints;
  uint64_t setup_serial = 0, geom_serial = 0;
  struct PressureSetup {
    bool valid = false;
    int method = -1;
    int A[7] = {};
    uint64_t setup_serial = 0, geom_serial = 0;
    bool classes = false, shadow = false;
    bool ref_active = false;
    double ref[3] = {0.0, 0.0, 0.0};
    std::vector<int> active;
    ffvhc::BiOpts o;
  } psetup;
  // stencil classes of the compact-record methods (FFVHC_OPTION_CLASSES): per slot the
  // class-uniform flag and the 27 class rows as FP32 bands and as 16-byte preconditioner records
  int* dcuni = nullptr;
  float* dctab = nullptr;
  void* dctrec = nullptr;
  bool classes = true;
  // shadow residual of the compact-record methods: r0 = r (default, the reference's choice) or the
  // seeded hash vector (FFVHC_OPTION_SHADOW 1: on the 1,128-block Poisson benchmark 2-4 fewer iterations for methods 24-27 and no
  // r0 reads, on the flow past a sphere 15% more iterations a step)
  bool shadow_hash = false;
  // replay the BiCGSTAB iteration as a captured HIP graph (ffvhc_ctx_set_option; off by default:
  // measured 1-11% slower per iteration than direct launches on ROCm 7.2, 64-960 blocks)
  bool graphs = false;
  std::vector<ffvhc::FieldRec> fields;
  long long bytes = 0;
  // free slot intervals [begin, end)
  std::vector<std::pair<int64_t, int64_t>> free_slots;
};

namespace ffvhc {
int refresh_active(ffvhc_ctx* ctx);
std::vector<int> selected_cases(ffvhc_ctx* ctx);
int ensure_second_buffer(ffvhc_ctx* ctx, int field);
int alloc_field(ffvhc_ctx* ctx, int prec, int* field);
int scratch_field(ffvhc_ctx* ctx, int prec, int role, int* field);
bool field_ok(ffvhc_ctx* ctx, int f, int prec);
int upload_case_meta(ffvhc_ctx* c, int id);
int host_impose_bc(ffvhc_ctx* c, int id, double* padded, int g, int bcg, int updater);
}  // namespace ffvhc
