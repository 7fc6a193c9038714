// Host-side launchers of the HIP kernels (all asynchronous on the given stream).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace ffvhc {

// scalar steps applied per case by launch_finalize after a reducing kernel
enum FinalOp {
  FIN_BI_INIT = 0,
  FIN_BI_ALPHA,
  FIN_BI_GAMMA,
  FIN_BI_XR,
  FIN_CG_INIT,
  FIN_CG_ALPHA,
  FIN_CG_XR,
  FIN_RBGS,
  FIN_DEFECT,
  FIN_FIXAVG,
  FIN_BI_REANCHOR
};
// gate: for FIN_BI_REANCHOR, the outer state; cases whose outer solve is done are skipped
void launch_finalize(hipStream_t sm, const int* caseact, int ncase, const DevCase* cs, SolveState* st,
                     const double* part, size_t nslots, int op, int arg, const SolveState* gate = nullptr);

template <class T>
void launch_ghost_pull(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                       const T* x, T* h, int bcg, int updater, bool multilevel);

template <class T>
void launch_rbgs_iter(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                      SolveState* st, Coef7<T> A, const T* b, const T* xin, T* xout, const T* hin, T* hout,
                      double* part, size_t nslots, int n, int bcg, int mode = 0);
template <class T>
void launch_shift_push(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                       const SolveState* st, T* x, const T* hin, T* hout, int bcg);

template <class T>
void launch_select(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const SolveState* st, T* x0,
                   const T* x1, T* h0, const T* h1);

template <class T>
void launch_fill(hipStream_t sm, const int* active, int nact, T* x, T* h, T v);
template <class T>
void launch_copy(hipStream_t sm, const int* active, int nact, T* y, T* hy, const T* x, const T* hx);

template <class T>
void launch_cg_init(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs, SolveState* st,
                    Coef7<T> A, const T* b, const T* x, const T* hx, T* r, double* part, size_t nslots);
template <class T>
void launch_cg_p(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                 const SolveState* st, Coef7<T> A, const T* r, T* p, T* hp, int count, int bcg);
template <class T>
void launch_cg_q(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs, SolveState* st,
                 Coef7<T> A, const T* p, const T* hp, T* q, double* part);
template <class T>
void launch_cg_xr(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs, SolveState* st,
                  Coef7<T> A, T* x, T* r, const T* p, const T* q, double* part, size_t nslots, int count);
template <class T>
void launch_bicgstab_init(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                          SolveState* st, Coef7<T> A, const T* b, const T* x, const T* hx, BiVecs<T> V, double* part,
                          size_t nslots, bool hash_r0 = false, double* bsr = nullptr);
// Precision options of an FP64 BiCGSTAB: A32 = FP32 copy of the matrix for the matvecs and the
// Jacobi preconditioner (widened to FP64 on load); gs = block red-black GS preconditioner with
// an FP32 matrix and FP32 arithmetic, an FP32 matrix and FP64 arithmetic, all FP64, or the
// symmetric (red, black, red) sweep with an FP32 matrix and FP64 arithmetic.
struct BiOpts {
  // GSC: `sweeps` (1-6) red-black SOR sweeps on the compact per-cell records of launch_gs_pack
  enum Gs { NONE = 0, GS32, GS32_ARITH64, GS64, SGS32_ARITH64, GSC };
  const Coef7<float>* A32 = nullptr;
  int gs = NONE;
  const Coef7<float>* M32 = nullptr;
  const Coef7<double>* M64 = nullptr;
  // shadow residual r0 = the seeded hash vector computed in the kernels that read it (nothing
  // stored or read for r0; the matvec and XR hashed variants exist for the A32 path)
  bool hash_r0 = false;
  const uint4* rec = nullptr;
  int sweeps = 1;
  // coarse correction (GSC only): FP32 inverse of A_c per case at offsets coff[case] (row pitch
  // n rounded up to a multiple of 4), the block sums of v in part[4 * nslots + slot], corrections
  // ce[mode * nslots + slot]; mult: applied before the sweeps (multiplicative), from the block
  // sums of r, q_, p in ce[2..4][slot]
  const float* cinv = nullptr;
  const long long* coff = nullptr;
  double* ce = nullptr;
  bool mult = false;
  // stencil classes (GSC only): class rows of every block as FP32 bands tab[slot][27][8] and as
  // records trec[slot][27], uni[slot] = 1 for a class-uniform block; nreg / nirr count the active
  // class-uniform blocks and the others (trec null: every block uses the per-cell arrays)
  const float* tab = nullptr;
  const uint4* trec = nullptr;
  const int* uni = nullptr;
  int nreg = 0, nirr = 0;
};

template <class T>
void launch_bicgstab_iter(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase,
                          const DevBlock* blk, const DevCase* cs, SolveState* st, Coef7<T> A, T* x, BiVecs<T> V,
                          double* part, size_t nslots, int count, int mask = 31, const BiOpts& o = BiOpts());
template <class T>
void launch_bicgstab_reanchor(hipStream_t sm, const int* active, int nact, const DevBlock* blk,
                              const SolveState* gate, BiVecs<T> V, double* part, size_t nslots);
template <class T>
void launch_bicgstab_shadow(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                            const SolveState* gate, BiVecs<T> V);
template <class T>
void launch_bicgstab_restart(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const SolveState* st,
                             BiVecs<T> V);
void launch_defect(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                   SolveState* st, Coef7<double> A, const double* b, const double* x, const double* hx, float* r32,
                   double* part, size_t nslots, int outer);
void launch_add_correction(hipStream_t sm, const int* active, int nact, const DevBlock* blk, SolveState* st,
                           const SolveState* inner, double* x, const float* e);
void launch_convert(hipStream_t sm, const int* active, int nact, float* y, const double* x);
// Compact GS preconditioner records (16 bytes per cell: 1 / Ap in FP32, A_q / Ap in FP16)
void launch_gs_pack(hipStream_t sm, const int* active, int nact, Coef7<double> A, uint4* rec);
// Stencil position classes of each block (uniform[slot] = 1 when the block is class-uniform) and
// its 27 class rows: FP32 bands tab[slot][27][8] and preconditioner records trec[slot][27]
void launch_stencil_classes(hipStream_t sm, const int* active, int nact, Coef7<double> A, int* uniform, float* tab,
                            uint4* trec);
// coarse operator rows (25 doubles per slot), dense Gauss-Jordan inversion, FP32 copy, apply
void launch_coarse_rows(hipStream_t sm, const int* active, int nact, const DevBlock* blk, Coef7<double> A,
                        double* rows);
void launch_coarse_assemble(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                            const long long* off, const double* rows, double* M, size_t total);
void launch_coarse_decouple(hipStream_t sm, double* M, int n);
// T: scratch of 32 * n doubles
void gauss_jordan_inverse(hipStream_t sm, double* M, int n, double* T);
void launch_phase_pass(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const uint8_t* cid8,
                       size_t capnc, int target, uint16_t* lab, int* changed);
void launch_phase_store(hipStream_t sm, const int* active, int nact, const uint16_t* lab, uint32_t* flag,
                        int* nfluid, int* irregular);
void launch_gjw_get_diag(hipStream_t sm, const double* M, int n, int k0, int kb, double* D);
void launch_gjw_get_cols(hipStream_t sm, const double* M, int n, int k0, int kb, double* C);
void launch_gjw_zero_tcols(hipStream_t sm, double* T, int n, int k0, int kb);
void launch_gjw_put_rows(hipStream_t sm, double* M, int n, int k0, int kb, const double* T, const double* D);
// FP32 copy of an n x n FP64 matrix with the row pitch np (>= n, pad zero)
void launch_to_f32(hipStream_t sm, const double* x, float* y, int n, int np);
void launch_coarse_apply(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                         const SolveState* st, const float* inv, const long long* off, const double* c, double* e);
// mode 1 also forms alpha from the Q partials apart (and stores it in the case state); mode 0
// with xpart also forms and stores the finalisation of iteration count - 1 from its XR partials
void launch_coarse_apply_mul(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                             SolveState* st, const float* inv, const long long* off, double* ce, size_t nslots,
                             int mode, const double* apart, int count = 0, const double* xpart = nullptr);

void launch_predict(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                    const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc);
void launch_abd_u(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                  const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc, int dir, double alpha);
void launch_p1(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
               const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc);
void launch_ab_p(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                 const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc);
void launch_refval(hipStream_t sm, const int* slots, int n, const DevBlock* blk, const StepFields& F, double xr,
                   double yr, double zr, double pr);
void launch_matrix_bc(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                      const StepFields& F, int bcg);
void launch_force_p(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase, const DevBlock* blk,
                    const DevCase* cs, const double* p, const double* hp, const uint8_t* cid8, const uint32_t* flag,
                    size_t capnc, int target, double* part, size_t nslots, double* out);
void launch_force_v(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase, const DevBlock* blk,
                    const DevCase* cs, const double* const* u, const double* const* hu, const float* cut,
                    const uint8_t* cid8, const uint32_t* flag, size_t capnc, int target, double* part, size_t nslots,
                    double* out);
void launch_stats(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase, const DevCase* cs,
                  const double* x, double* part, size_t nslots, double* out);
void launch_corr(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                 const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc);

// outer (optional): inner solve of a defect-correction / restart loop: cases done there are
// skipped, and each case's countMax is reduced by the iterations its outer state has used
void launch_solve_init(hipStream_t sm, SolveState* st, int ncase, double omega, double eps, int countMax,
                       const SolveState* outer = nullptr);
// outer[c].used += inner[c].count
void launch_add_used(hipStream_t sm, SolveState* outer, const SolveState* inner, int ncase);

}  // namespace ffvhc
