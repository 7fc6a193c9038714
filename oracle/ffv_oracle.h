/*
 * ffv_oracle.h -- CPU restatement of FFV-HC's fractional-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (ffv-hc_amd/) includes,
 * links or calls this code.  It is the checker the HIP path is compared with:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * Layout follows the reference exactly: one padded block of
 * (nx+2g)*(ny+2g)*(nz+2g) reals per field, x fastest, Fortran index
 * i in [1-g, nx+g] (src/CGMfuncs/bils.f90:17, Solver.cpp:588).
 *
 * Pinning (see tests/test_oracle_ref.py, tests/test_oracle_math.py):
 *  - every per-block kernel below is compared bit-for-bit with the reference
 *    Fortran compiled all-double (oracle/Makefile, oracle/_ref/libffvref.so)
 *    and within 1e-13 (relative to the field's scale) with the reference's
 *    as-configured build (-fdefault-real-8, quad literals; libffvref_q.so);
 *  - the Updater5 gradient prolongation / restriction (a C++ template of the
 *    reference, not buildable here) is pinned by closed forms: linear fields
 *    are reproduced, constants are preserved, hand-computed cases;
 *  - grid-level orchestration (ghost fill order, RBGS/CG/fixed-average
 *    control flow, the time step) is pinned by invariants and closed forms
 *    (textbook red-black GS on one block, scipy CG on an SPD system, manufactured
 *    Poisson solutions, uniform-flow preservation, divergence after projection).
 *  - BCMTools ghost scheduling order (AtOnce) is not vendored: the oracle uses
 *    F2C -> same level -> C2F -> outer BC (LocalScalar3D.h:252-262 order);
 *    parity of that choice is unpinned against the reference.
 */
#ifndef FFV_ORACLE_H
#define FFV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double real;

/* ---------------- per-block kernels (reference Fortran signatures) -------- */
/* bils.f90 */
void orc_jacobi_smoother(real* x1, const real* x0, const real* Ap, const real* Aw, const real* Ae,
                         const real* As, const real* An, const real* Ab, const real* At, const real* b,
                         real omega, const int* sz, int g);
void orc_rbgs_smoother(real* x, const real* Ap, const real* Aw, const real* Ae, const real* As,
                       const real* An, const real* Ab, const real* At, const real* b, real omega,
                       int color, int offset, const int* sz, int g);
void orc_calc_ax(real* Ax, const real* Ap, const real* Aw, const real* Ae, const real* As,
                 const real* An, const real* Ab, const real* At, const real* x, const int* sz, int g);
void orc_calc_r(real* r, const real* Ap, const real* Aw, const real* Ae, const real* As,
                const real* An, const real* Ab, const real* At, const real* x, const real* b,
                const int* sz, int g);
real orc_calc_r2(const real* Ap, const real* Aw, const real* Ae, const real* As, const real* An,
                 const real* Ab, const real* At, const real* x, const real* b, const int* sz, int g);
/* blas.f90 */
void orc_fill(real* x, real a, const int* sz, int g);
void orc_copy(real* y, const real* x, const int* sz, int g);
void orc_adda(real* y, real a, const int* sz, int g);
void orc_axpy(real* y, const real* x, real a, const int* sz, int g);
void orc_xpay(real* y, const real* x, real a, const int* sz, int g);
void orc_axpyz(real* z, const real* x, const real* y, real a, const int* sz, int g);
void orc_axpbypz(real* z, const real* x, const real* y, real a, real b, const int* sz, int g);
real orc_dot(const real* x, const real* y, const int* sz, int g);
/* bsf3d.f90 */
real orc_sf3d_calc_sum(const real* x, const int* sz, int g);
/* sup.f90 */
void orc_sup_copy_from_neighbor(real* dst, const int* i1_dst, const real* src, const int* i1_src,
                                const int* sz_c, const int* sz, int g);
void orc_sup_copy_from_neighbor_c2f(real* dst, const int* i1_dst, const real* src, const int* i1_src,
                                    const int* sz_c, const int* sz, int g);
void orc_sup_copy_from_neighbor_f2c(real* dst, const int* i1_dst, const real* src, const int* i1_src,
                                    const int* sz_c, const int* sz, int g);
/* Scalar3DUpdater5.h (velocity ghosts): face = 0..5, subface = 0..3 */
void orc_u5_f2c(real* coarse, const real* fine, int face, int subface, const int* sz, int g);
void orc_u5_c2f(real* fine, const real* coarse, int face, int subface, const int* sz, int g);
/* bcx.f90: face = 0..5 (X_M..Z_P), type 0 Dirichlet, 1 Neumann, 9 periodic; xc already
 * scaled (Neumann value * dx, BlockScalar3D.h:74-76) */
void orc_bc_x(real* x, int face, int type, real xc, const int* sz, int g);
/* bca.f90 */
void orc_bc_a(real* Ap, real* Anb, real* b, int face, int type, real xc, const int* sz, int g);
/* bcut.f90 (hot subset) */
void orc_bcut_calc_c_f_c2(real* fc, const real* f, const real* vw, const real* ve, const real* vs,
                          const real* vn, const real* vb, const real* vt, const real* const* c,
                          const int* const* cid, const int* pid, real dx, real dt, real fi,
                          const int* sz, int g);
void orc_bcut_calc_d_u(real* ud, const real* u, const real* const* c, const int* const* cid,
                       const int* pid, real rho, real mu, real dx, real dt, real Uc, const int* sz, int g);
void orc_bcut_update_u(real* u, const real* uc, const real* ucp, const real* ud, const real* udp,
                       const real* p, const real* const* c, const int* const* cid, const int* pid,
                       int dir, real rhof, real mu, real dx, real dt, real Uc, real gx, real gy, real gz,
                       const int* sz, int g);
void orc_bcut_calc_abd_u(real* const* A, real* b, real* u, const real* uc, const real* ucp, real* ud,
                         const real* p, const real* const* c, const int* const* cid, const int* pid,
                         int dir, real alpha, real rhof, real mu, real dx, real dt, real Uc,
                         real gx, real gy, real gz, const int* sz, int g);
void orc_bcut_add_g(real* ux, real* uy, real* uz, const real* t, real gx, real gy, real gz, real betag,
                    real tr, real dx, real dt, const int* sz, int g);
void orc_bcut_add_f(real* ux, real* uy, real* uz, const real* fx, const real* fy, const real* fz,
                    const int* pid, real rhof, real dx, real dt, const int* sz, int g);
void orc_bcut_remove_p(real* ux, real* uy, real* uz, const real* p, const real* const* c,
                       const int* const* cid, const int* pid, real rhof, real dx, real dt,
                       const int* sz, int g);
void orc_bcut_calc_ab_p(real* const* A, real* b, real* const* v, const real* ux, const real* uy,
                        const real* uz, const real* const* c, const int* const* cid, const int* pid,
                        real rhof, real dx, real dt, const int* sz, int g);
void orc_bcut_corr_u(real* ux, real* uy, real* uz, real* const* v, real* lapp, const real* p,
                     const real* const* c, const int* const* cid, const int* pid, real rhof, real dx,
                     real dt, const int* sz, int g);
void orc_bcut_set_reference_value(real* const* A, real* b, real xr, real yr, real zr, real pr, real dx,
                                  const real* org, const int* sz, int g);
void orc_bcut_calc_f_p(real* fspx, real* fspy, real* fspz, real* fsp, int cid_target, const real* p,
                       const real* const* c, const int* const* cid, const int* pid, real dx, real dt,
                       const int* sz, int g);
void orc_bcut_calc_f_v(real* fsvx, real* fsvy, real* fsvz, real* fsv, int cid_target, const real* ux,
                       const real* uy, const real* uz, const real* const* c, const int* const* cid,
                       const int* pid, real rho, real mu, real dx, real dt, real Uc, const int* sz, int g);
/* bstl.f90 (setup) */
void orc_bstl_cutoff(real* const* c, int* const* cid, real eps, const int* sz, int g);
int orc_bstl_detect_zerocut(real* const* c, int* const* cid, const int* sz, int g);
int orc_bstl_fill_holes_v1(real* const* c, int* const* cid, int bClose, const int* sz, int g);
int orc_bstl_fill_holes_v2(real* const* c, int* const* cid, int bClose, const int* sz, int g);

/* ---------------- grid level (many blocks) -------------------------------- */
/* Block table shared with the product (built by the product's octree code):
 *   ftype[b*6+f]: 0 outer face, 1 same-level neighbour, 2 neighbour coarser
 *                 (this block receives C2F), 3 neighbours finer (receives F2C);
 *   nbr[(b*6+f)*4+s]: neighbour block (s = subface for ftype 3, s = 0 otherwise);
 *   sub[b*6+f]: for ftype 2, this block's subface on the coarse neighbour's face.
 * Subface s has tangential origins (s & 1, s >> 1) on the axes (y,z) of X faces,
 * (z,x) of Y faces and (x,y) of Z faces (Scalar3DUpdater1.cpp:264-343). */
typedef struct {
  int nblk;
  int n;  /* cells per block edge */
  int g;  /* ghost layers */
  const int* ftype;
  const int* nbr;
  const int* sub;
  const double* dx;  /* cell size per block */
  const double* org; /* block origin [nblk*3] */
} orc_grid;

/* LocalScalar3D::ImposeBoundaryCondition (LocalScalar3D.h:266-296): VC update then outer BC.
 * updater 1 = Scalar3DUpdater1 (P, default fields), 5 = Scalar3DUpdater5 (U).
 * bctype/bcval [6]: outer-face BC (type -1 = none). */
void orc_ghost_fill(const orc_grid* gr, real* f, int updater, const int* bctype, const real* bcval);
/* LocalScalar3D::ImposeBoundaryCondition(A..., b) (LocalScalar3D.h:298-323) */
void orc_matrix_bc(const orc_grid* gr, real* const* A, real* b, const int* bctype, const real* bcval);
real orc_grid_dot(const orc_grid* gr, const real* x, const real* y);
real orc_grid_r2(const orc_grid* gr, real* const* A, const real* x, const real* b);
real orc_grid_sum(const orc_grid* gr, const real* x);

/* FFVILS (FFVILS.h) -- return 0, or 1 if the residual became NaN/Inf (reference Exit(0)) */
int orc_rbgs(const orc_grid* gr, real* x, real* const* A, const real* b, const int* bctype,
             const real* bcval, real omega, int countMax, real eps, int* count, real* residual);
int orc_rbgs_fixed_average(const orc_grid* gr, real* x, real* const* A, const real* b,
                           const int* bctype, const real* bcval, int Ncells, real omega, int countMax,
                           real eps, int* count, real* residual);
int orc_jacobi(const orc_grid* gr, real* x, real* const* A, const real* b, real* x0, const int* bctype,
               const real* bcval, real omega, int countMax, real eps, int* count, real* residual);
/* FFVILS::CG (FFVILS.h:661-791); scratch r,p,q,z,x0 carry the field's BC (bctype) for ghosts */
int orc_cg(const orc_grid* gr, real* x, real* const* A, const real* b, real* r, real* p, real* q, real* z,
           real* x0, const int* bctype, const real* bcval, real omega, int countPre, int countMax,
           real eps, int* count, real* residual);

/* FFVILS::BiCGSTAB (FFVILS.h:793-974); sc = {r, r0, p, p_, q_, s, s_, t_, x0} */
int orc_bicgstab(const orc_grid* gr, real* x, real* const* A, const real* b, real* const* sc, const int* bctype,
                 const real* bcval, real omega, int countPre, int countMax, real eps, int* count, real* residual);

/* Fractional step (Solver::Update, Solver.cpp:3344-3416) with explicit U (UpdateUXe 3535-3744),
 * UpdateP (5294-5688) and UpdateU (5690-5771).  Temperature/fan stages are out of scope. */
typedef struct {
  real *ux, *uy, *uz, *p;
  real *uc[3], *ucp[3], *ud[3], *udp[3];
  real *v[6];       /* vw ve vs vn vb vt */
  real *A[7], *b;   /* Ap Aw Ae As An Ab At */
  real *lapp;
  real *t0, *f[3];  /* temperature (buoyancy) and body force, executed as in UpdateP */
  real *cgs[5];     /* CG scratch r p q z x0 */
  real *bis[9];     /* BiCGSTAB scratch r r0 p p_ q_ s s_ t_ x0 (implicit U) */
  real *c[6];
  int *cid[6];
  int *pid;
} orc_fields;

typedef struct {
  double rho, mu, dt, gx, gy, gz, betag, tr;
  int bct_u[3][6];
  double bcv_u[3][6];
  int bct_p[6];
  double bcv_p[6];
  int solver; /* 0 RBGS, 1 RBGS_WithFixedAverage, 2 CG (Jacobi PCG), 3 Jacobi-BiCGSTAB */
  double omega;
  int countMax;
  double eps;
  int countPre;
  int ncells; /* numTotalCells (fixed-average mean) */
  int accel_steps; /* TimeControl/Acceleration (0 = off): ramps ux X_M inflow */
  int ref_active; double ref_x, ref_y, ref_z, ref_p;
  /* implicit U (UpdateUX/UY/UZ, Solver.cpp:4391-4690): calc_abd_u + BiCGSTAB per component */
  int u_implicit;
  double alpha_u, omega_u, eps_u;
  int countMax_u;
} orc_params;

typedef struct {
  int countP;
  double residualP;
  int status;
  double times[7];
  int countU[3];
  double residualU[3];
} orc_step_info;

void orc_update(const orc_grid* gr, orc_fields* F, const orc_params* P, int step, orc_step_info* info);
void orc_set_threads(int n);
/* Planted mistakes for the pin tests (0 = none): 1 BiCGSTAB beta sign, 2 BiCGSTAB alpha/gamma
 * swapped in the x update, 3 Jacobi driver drops each sweep's result, 4 RBGS_WithFixedAverage
 * without the mean shift, 5 UpdateU before the pressure solve, 6 bcut_remove_p skipped,
 * 7 CG beta inverted, 8 ghost fill with C2F before F2C. */
void orc_plant(int id);

#ifdef __cplusplus
}
#endif
#endif
