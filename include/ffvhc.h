/*
 * ffvhc.h -- C ABI of the MI355X-native FFV-HC fractional-step hot path.
 *
 * Everything is plain C: opaque handles, plain pointers and sizes, int status codes.
 * Device memory is owned by the context; hosts hold handles.  Status codes:
 *   0 ok, -1 invalid argument, -2 HIP error, -3 out of capacity, -4 not found,
 *   1 NaN/Inf residual (the reference calls Exit(0) there, FFVILS.h:36-43),
 *   2 CFL guard tripped (Solver.cpp:2665-2675).
 * ffvhc_last_error() describes the last non-zero status of the calling thread.
 *
 * Field arrays crossing the boundary use the reference's per-block layout: for each block in
 * block order, a padded (n+2g)^3 array, x fastest (LocalScalar3D.h:232-235, Solver.cpp:588).
 * The device keeps the interior and the first ghost layer (the only one a 7-point stencil
 * reads); download fills those cells and leaves the rest of the caller's array untouched.
 */
#ifndef FFVHC_H
#define FFVHC_H

#ifdef __cplusplus
extern "C" {
#endif

#define FFVHC_MAX_REGIONS 32

/* ------------------------------------------------------------------ block octree (host) */

/* Refinement region; kind 0 ball, 1 spherical shell (SphereDivider2.h:75-155), 2 box
 * (BoundingBox lists, SphereDivider3.h:72-112). */
typedef struct ffvhc_region {
  int kind;
  double a[3];
  double b[3];
  double r;
  double dr;
  int level;
} ffvhc_region;

/* DomainInfo/RootBlock + Tree of the reference configuration (FFVConfig.cpp:36-127). */
typedef struct ffvhc_tree_spec {
  int root[3];          /* RootBlock/Grid */
  double origin[3];     /* RootBlock/Origin */
  double root_length;   /* RootBlock/Length */
  int periodic[3];      /* RootBlock/PeriodicX..Z */
  int n;                /* LeafBlock/NumberOfCells (16 on the device) */
  int min_level;        /* Tree/MinLevel */
  double margin;        /* search-region margin, in cells of the finest level (NumberOfMarginalCells) */
  int ordering;         /* Tuning/BlockOrdering: 0 Z, 1 Hilbert */
  int nregions;
  ffvhc_region regions[FFVHC_MAX_REGIONS];
} ffvhc_tree_spec;

typedef struct ffvhc_tree ffvhc_tree;

/* Replaces BCMOctree construction + Partition + makeNeighborInfo (Solver.cpp:306-387). */
int ffvhc_tree_build(const ffvhc_tree_spec* spec, ffvhc_tree** out);
int ffvhc_tree_num_blocks(const ffvhc_tree* tree);
int ffvhc_tree_max_level(const ffvhc_tree* tree);
/* Block table, block order: level[nb], org[3nb], dx[nb], ftype[6nb] (0 outer, 1 same level,
 * 2 coarser neighbour, 3 finer neighbours), nbr[24nb], sub[6nb]; any pointer may be NULL. */
int ffvhc_tree_table(const ffvhc_tree* tree, int* level, double* org, double* dx, int* ftype, int* nbr, int* sub);
void ffvhc_tree_free(ffvhc_tree* tree);

/* ------------------------------------------------------------------ geometry (host) */

/* Sphere as cut cells: replaces Cutlib's CalcCutInfo (Solver.cpp:530-682, cut positions stored
 * as float32) with an analytic sphere, then ModifyCut1/2 (bstl_cutoff, zero-cut detection,
 * hole filling v1/v2/v3, Solver.cpp:684-1046) and the phase flood fill from the seed point
 * (FillCellFlag/InitPhase, Solver.cpp:1182-1287, 1404-1537). */
typedef struct ffvhc_sphere_spec {
  double center[3];
  double radius;
  int cut_id;           /* boundary id of the sphere surface (a phase boundary) */
  double cutoff;        /* ShapeApproximation/Cutoff */
  int hole_filling[3];  /* GridGeneration/HoleFilling, HoleFilling2, HoleFilling3 */
  double seed[3];       /* ApplicationControl/Filling/Origin (a fluid point) */
} ffvhc_sphere_spec;

/* Outputs per block, padded with g ghost layers: c[6][nb][npad] (real), cid[6][nb][npad],
 * pid[nb][npad] with npad = (n+2g)^3.  counts = {fluid cells, solid cells, hole faces v1,
 * hole faces v2}.  Any output pointer may be NULL except pid. */
int ffvhc_geometry_sphere(const ffvhc_tree* tree, const ffvhc_sphere_spec* sphere, int g, double* c, int* cid,
                          int* pid, long long* counts);

/* ------------------------------------------------------------------ device context */

typedef struct ffvhc_ctx ffvhc_ctx;

const char* ffvhc_last_error(void);

/* One context per GPU; max_blocks bounds the blocks of all resident cases (every field is
 * allocated for max_blocks blocks of 16^3 cells). */
int ffvhc_ctx_create(int device, long long max_blocks, ffvhc_ctx** out);
void ffvhc_ctx_destroy(ffvhc_ctx* ctx);
int ffvhc_ctx_sync(ffvhc_ctx* ctx);
/* bytes of device memory the context holds */
long long ffvhc_ctx_device_bytes(const ffvhc_ctx* ctx);
/* Context options (no counterpart in the reference).  FFVHC_OPTION_GRAPHS: 1 replays the
 * BiCGSTAB iteration as a HIP graph captured once per solve, 0 (default) launches its kernels
 * one by one; the results are identical (the graph replay measured slower on ROCm 7.2). */
#define FFVHC_OPTION_GRAPHS 1
/* FFVHC_OPTION_CLASSES: 1 (default) lets the compact-record methods (ffvhc_solve_mixed 11-28)
 * find the blocks whose seven bands are the same within each of the 27 stencil position classes
 * (face layers, edges, corners, inside: every block without cut cells) and read those blocks'
 * bands and preconditioner records from 27 rows instead of per cell; 0 reads every block per
 * cell.  The results are identical. */
#define FFVHC_OPTION_CLASSES 2
/* FFVHC_OPTION_SHADOW: 0 (default) the compact-record methods (11-28) take the reference's shadow
 * residual r0 = r; 1 the seeded hash vector of method 10, computed where it is read (no r0
 * vector is stored or read).  The iteration count changes (on the 1,128-block Poisson benchmark 2-4 fewer,
 * on the flow past a sphere more); the accepted answer meets the same FP64 residual. */
#define FFVHC_OPTION_SHADOW 3
/* FFVHC_OPTION_SETUP_REUSE: 1 (default) lets ffvhc_step keep the setup of its mixed-precision
 * pressure solve (methods 11-28: FP32 band copies, preconditioner records, stencil class tables,
 * coarse inverse) from one step to the next while the pressure matrix cannot have changed: the
 * step rebuilds A from the geometry, the density and the BC types, and the setup is made again
 * after a case is added or removed, a BC is set, the active blocks, the matrix fields, the
 * method or an option change, or any other solve writes the setup buffers.  0 makes the setup
 * every step.  The results are identical. */
#define FFVHC_OPTION_SETUP_REUSE 4
/* FFVHC_OPTION_COARSE_GEMM: the smallest case (in blocks) whose coarse operator (two-level methods
 * 17-28) is inverted with 256-wide panels whose products run as FP64 GEMMs (matrix cores); smaller
 * ones use 32-wide panels on the vector units.  Default 2048 (measured: at 1,128 blocks the
 * vector form is as fast, at 4,048 the GEMM form takes 10 ms instead of 22, at 8,512 41 instead of
 * 182).  Both are Gauss-Jordan without pivoting; their inverses differ in rounding only. */
#define FFVHC_OPTION_COARSE_GEMM 5
/* FFVHC_OPTION_FUSE_FIN: 1 (default) lets the multiplicative two-level methods (23-28) form the
 * end of each BiCGSTAB iteration (convergence test, next rho and beta) inside the next iteration's
 * first coarse application instead of a launch of its own, except after the last iteration of a
 * launch batch; 0 launches it every iteration.  The results are identical. */
#define FFVHC_OPTION_FUSE_FIN 6
int ffvhc_ctx_set_option(ffvhc_ctx* ctx, int option, int value);

/* Physical parameters and outer BCs of one flow case.  BC groups: 0 UX, 1 UY, 2 UZ, 3 P,
 * 4 Neumann-0 fields (boundaryTypeNULL, Solver.cpp:2024-2034), 5 zero Dirichlet.  Types
 * -1 none, 0 Dirichlet, 1 Neumann (value per unit length, scaled by dx as BlockScalar3D.h:74-76),
 * 9 periodic. */
typedef struct ffvhc_case_params {
  double rho, mu, dt;
  double gx, gy, gz, betag, tr;
  int bct[6][6];
  double bcv[6][6];
} ffvhc_case_params;

/* Add a case: its blocks get device slots; geometry arrays as from ffvhc_geometry_sphere
 * (g ghost layers) or NULL for an all-fluid, cut-free grid. */
int ffvhc_case_add(ffvhc_ctx* ctx, const ffvhc_tree* tree, const ffvhc_case_params* params, int g, const double* c,
                   const int* cid, const int* pid, int* case_id);
int ffvhc_case_remove(ffvhc_ctx* ctx, int case_id);
int ffvhc_case_num_blocks(const ffvhc_ctx* ctx, int case_id);
/* LocalScalar3D::ResetBoundaryConditionValue (LocalScalar3D.h:325-335) */
int ffvhc_case_set_bc(ffvhc_ctx* ctx, int case_id, int group, int face, int type, double value);
/* Phase flood fill on the device (FillCellFlag / InitPhase, Solver.cpp:1182-1287, 1404-1537): the
 * case's cells reachable from the cell holding `seed` through faces whose cut id is not cut_id
 * become fluid, all others solid (a move is blocked when the entered cell's face toward the cell
 * it comes from carries cut_id), from the cut ids the case was added with.  counts[0..1] = fluid,
 * solid cells (may be null).  Status -1 when the seed lies outside the case's blocks. */
int ffvhc_case_fill_phase(ffvhc_ctx* ctx, int case_id, const double* seed, int cut_id, long long* counts);
/* The case's phase (1 fluid, 0 solid) as padded int blocks of ghost width g (ghosts 0). */
int ffvhc_case_download_phase(ffvhc_ctx* ctx, int case_id, int g, int* pid);
/* Restrict subsequent calls to these cases (NULL / 0 = every resident case). */
int ffvhc_ctx_select(ffvhc_ctx* ctx, const int* case_ids, int ncases);

/* ------------------------------------------------------------------ fields */

/* Field storage for every slot: precision 8 (FP64) or 4 (FP32); nbuf 1 or 2 (ping-pong). */
int ffvhc_field_alloc(ffvhc_ctx* ctx, int precision, int* field);
int ffvhc_field_free(ffvhc_ctx* ctx, int field);
int ffvhc_field_upload(ffvhc_ctx* ctx, int field, int case_id, int g, const void* padded);
int ffvhc_field_download(ffvhc_ctx* ctx, int field, int case_id, int g, void* padded);
/* Checkpoint in the reference's Dump2 layout (LocalScalar3D.cpp:215-264): six ints {nx, ny, nz,
 * vc, nE = 1, nB} then each block's padded (n+2vc)^3 FP64 array; g = vc (1 or 2).  As Dump2, the
 * ghost layers are filled first (ImposeBoundaryCondition with BC group bc_group and updater 1 or
 * 5: fine-to-coarse, same level, coarse-to-fine, then the outer BC over the whole padded face);
 * cells none of these writes (block edges and corners away from outer faces) are written as 0,
 * where the reference writes whatever they last held.  The file name is the caller's (the
 * reference: <path>/<prefix><label>-<rank:5>-<step:10>.bin). */
int ffvhc_field_dump(ffvhc_ctx* ctx, int field, int case_id, int g, int bc_group, int updater, const char* path);
/* Load2 (LocalScalar3D.cpp:266-342): the same-resolution and the half-resolution (8^3 blocks,
 * sf3d_copy_x2) branches; then call ffvhc_ghost_fill. */
int ffvhc_field_load(ffvhc_ctx* ctx, int field, int case_id, const char* path);
int ffvhc_dump_header(const char* path, int* head6);
/* FFVILS::Fill (FFVILS.h:425-445): every cell and ghost = value */
int ffvhc_field_fill(ffvhc_ctx* ctx, int field, double value);
/* FFVILS::Copy (FFVILS.h:447-466) */
int ffvhc_field_copy(ffvhc_ctx* ctx, int dst, int src);

/* LocalScalar3D::ImposeBoundaryCondition (LocalScalar3D.h:266-296): ghost update then outer BC.
 * updater 1: Scalar3DUpdater1 (pressure, default fields); 5: Scalar3DUpdater5 (velocity). */
int ffvhc_ghost_fill(ffvhc_ctx* ctx, int field, int bc_group, int updater);

/* ------------------------------------------------------------------ linear solvers */

/* A = 7 fields {Ap, Aw, Ae, As, An, Ab, At}; counts/residuals are per selected case, in the
 * order of ffvhc_ctx_select (or case id order).  Return 1 if any case hit a NaN/Inf residual. */

/* FFVILS::RBGS (FFVILS.h:547-594) */
int ffvhc_rbgs(ffvhc_ctx* ctx, int x, const int* A, int b, int bc_group, double omega, int countMax, double eps,
               int* count, double* residual);
/* Exactly nsweep fused RBGS iterations (FFVILS::RBGS with eps < 0, countMax = nsweep) on the
 * context stream with HIP events around every launch; ms[0] = device time of the iteration
 * kernels, ms[1] = device time of the per-case residual finalisations (timing entry). */
int ffvhc_rbgs_fixed(ffvhc_ctx* ctx, int x, const int* A, int b, int bc_group, double omega, int nsweep, double* ms);
/* FFVILS::RBGS_WithFixedAverage (FFVILS.h:596-659): after every sweep the mean over the case's
 * cells (Ncells = blocks x 16^3) is removed, then the ghost layer is filled. */
int ffvhc_rbgs_fixed_average(ffvhc_ctx* ctx, int x, const int* A, int b, int bc_group, double omega, int countMax,
                             double eps, int* count, double* residual);
/* FFVILS::CG (FFVILS.h:661-791): Jacobi-preconditioned CG, countPre damped-Jacobi sweeps */
int ffvhc_cg(ffvhc_ctx* ctx, int x, const int* A, int b, int bc_group, double omega, int countPre, int countMax,
             double eps, int* count, double* residual);
/* FFVILS::BiCGSTAB (FFVILS.h:793-974): Jacobi right-preconditioned BiCGSTAB, countPre = 1.  The
 * coarse/fine ghost interpolation makes the octree Poisson operator non-symmetric; this is the
 * solver that converges on it (and the reference's default for P and implicit U). */
int ffvhc_bicgstab(ffvhc_ctx* ctx, int x, const int* A, int b, int bc_group, double omega, int countPre,
                   int countMax, double eps, int* count, double* residual);
/* niter BiCGSTAB iterations with HIP events between the launches; ms[0..4] = total device time of
 * P (p update + preconditioner), Q (q_ = A p_), S, T (t_ = A s_), XR; ms[5] = the per-case
 * scalar finalisations; ms[6] = the coarse corrections of the two-level methods (else 0).
 * ms holds 7 doubles (timing entry). */
int ffvhc_bicgstab_timing(ffvhc_ctx* ctx, int x, const int* A, int b, double omega, int niter, double* ms);
/* Coarse operator of the two-level methods (17-28) for one case, as the solver builds it:
 * ac = P^T A P (P = block indicator; n x n, n = the case's blocks, row-major, block order) and
 * ainv = its inverse from the device's blocked Gauss-Jordan.  No counterpart in the reference;
 * exported so the coarse operator can be checked against its definition. */
int ffvhc_coarse_operator(ffvhc_ctx* ctx, const int* A, int case_id, double* ac, double* ainv);
/* Stencil classes (no counterpart in the reference): of the selected cases' blocks, how many have
 * the same seven FP64 bands A throughout each of the 27 position classes (n_uniform; the compact-
 * record methods read their bands from 27 rows, see FFVHC_OPTION_CLASSES) and how many not. */
int ffvhc_stencil_classes(ffvhc_ctx* ctx, const int* A, int* n_uniform, int* n_other);
/* The same stage timing for the FP64-recurrence methods 5-28 of ffvhc_solve_mixed (FP64 x, A, b;
 * omega = GS relaxation factor): ms[0..4] = P, Q, S, T, XR kernels, ms[5] = finalisations,
 * ms[6] = coarse corrections (methods 17-29; 29: the inner solve's FP32 kernels on FP32 copies of
 * A and b from x = 0). */
int ffvhc_bicgstab_timing_mixed(ffvhc_ctx* ctx, int method, int x, const int* A, int b, double omega, int niter,
                                double* ms);
/* niter PCG iterations (no convergence test) with HIP events between the launches; ms[0..2] =
 * total device time of the p-update, matvec and x/r-update kernels, ms[3] = the per-case scalar
 * finalisations (timing entry). */
int ffvhc_cg_timing(ffvhc_ctx* ctx, int x, const int* A, int b, double omega, int niter, double* ms);
/* FFVILS::Jacobi (FFVILS.h:492-545) */
int ffvhc_jacobi(ffvhc_ctx* ctx, int x, const int* A, int b, int bc_group, double omega, int countMax, double eps,
                 int* count, double* residual);

/* True FP64 relative residual ||b - Ax||_2 / ||b||_2 of every selected case (FFVILS::CalcR2 and
 * IsConverged, FFVILS.h:247-296 and 24-47); x's ghost layer must be current. */
int ffvhc_residual(ffvhc_ctx* ctx, int x, const int* A, int b, double* residual);

/* Mixed precision: inner sweeps / matvecs in FP32 wrapped in FP64 defect correction to the same
 * relative residual eps (FP64 residual, dots and state).  method 0 RBGS, 1 Jacobi-PCG,
 * 2 Jacobi-BiCGSTAB (each restarted per correction); 3 Jacobi-BiCGSTAB with reliable updates (the
 * FP32 recurrence continues across corrections; inner_eps = residual reduction between updates);
 * 4 as 3 with a seeded random shadow residual r0 instead of r0 = r.
 * Methods 5-16 keep the BiCGSTAB recurrence, dots and state in FP64 and narrow what is stored:
 * 5 FP32 block red-black GS preconditioner (FP32 arithmetic), 6 the same GS all in FP64,
 * 7 FP32 copy of the matrix for the matvecs and the Jacobi preconditioner, 8 as 7 with the block
 * GS preconditioner in FP64 arithmetic, 9 as 8 with the symmetric (red, black, red) sweep, 10 as 8
 * with a seeded hash vector as the shadow residual r0 (computed where it is read, never stored),
 * 11-16 as 8 with 1-6 red-black SOR sweeps (block-local, from zero, FP64 arithmetic) on a compact
 * copy of the matrix: 1 / Ap in FP32 and A_q / Ap in FP16, 16 bytes per cell, 17-22 as 11-16
 * plus a coarse correction P A_c^-1 P^T (one value per block, A_c = P^T A P inverted densely),
 * 23-28 the same correction applied first and the sweeps on the remaining residual
 * (multiplicative two-level).
 * Method 29 is the defect-correction form with that preconditioner: each correction solves
 * A e = r (r the FP64 residual, rounded to FP32) by method 26's BiCGSTAB on FP32 vectors to the
 * relative tolerance inner_eps (at most countMax iterations), then x += e in FP64.
 * Methods 2, 3 and 4 are experimental, kept for comparison only: on every workload measured here
 * 2 is 4x slower than the FP64 Jacobi-BiCGSTAB and 3 / 4 stall near a relative residual of 1e-4
 * (the FP32 recurrence loses the bi-orthogonality coefficient); the autotuner never picks them.
 * For 5 and 7-28 the answer is accepted on the FP64 residual of the FP64 matrix; a case above
 * eps restarts from it (inner_eps unused; omega is the GS / SOR relaxation factor).
 * inner_eps: relative reduction per correction; countMax caps each inner solve (a capped inner
 * solve still yields a correction).  count[c] = inner iterations in total, outer[c] = corrections
 * taken, residual[c] = FP64 ||b - Ax|| / ||b|| after the last correction. */
int ffvhc_solve_mixed(ffvhc_ctx* ctx, int method, int x, const int* A, int b, int bc_group, double omega,
                      int countMax, double eps, double inner_eps, int* count, double* residual, int* outer);

/* ------------------------------------------------------------------ fractional time step */

/* Field handles of Solver::Update (Solver.cpp:3344-3416), all FP64.  uc/ud receive this step's
 * convection/diffusion terms and ucp/udp hold the previous step's: on return the library swaps
 * uc<->ucp and ud<->udp in the struct, which leaves ucp/udp holding this step's terms exactly as
 * the reference's copies do (Solver.cpp:3725-3732).  t and f may be -1 (zero temperature
 * deviation / zero body force; the additions are still evaluated, as the reference does). */
typedef struct ffvhc_step_fields {
  int ux, uy, uz, p;
  int uc[3], ucp[3], ud[3], udp[3];
  int v[6];    /* face velocities vw ve vs vn vb vt */
  int A[7], b; /* Poisson matrix Ap..At and right-hand side */
  int lapp;    /* Laplacian of p (bcut_corr_u) */
  int t;
  int f[3];
} ffvhc_step_fields;

/* SolvingMethod/P and its parameters (FFVConfig.cpp:164-165, 434) */
typedef struct ffvhc_step_params {
  int solver;        /* 0 RBGS, 1 RBGS_WithFixedAverage, 2 Jacobi-PCG, 3 Jacobi-BiCGSTAB */
  int mixed;         /* 1: FP32 inner solve + FP64 defect correction to eps */
  double omega;
  int countMax;
  double eps;
  int countPre;      /* must be 1 for PCG / BiCGSTAB */
  double inner_eps;  /* mixed only */
  int ref_active;    /* bcut_set_reference_value (off by default) */
  double ref[3];
  double ref_p;
  /* SolvingMethod/U: 0 explicit (UpdateUXe, Adams-Bashforth 2), 1 implicit (UpdateUX:
   * bcut_calc_abd_u + Jacobi-BiCGSTAB per component, Solver.cpp:4391-4690) */
  int u_implicit;
  double alpha_u;    /* TimeIntegrationMethodForU_Alpha */
  double omega_u;
  int countMax_u;
  double eps_u;
  /* with mixed = 1: the ffvhc_solve_mixed method for P (-1: by `solver`: RBGS 0, PCG 1,
   * BiCGSTAB 2); omega is then that method's relaxation factor */
  int mixed_method;
} ffvhc_step_params;

typedef struct ffvhc_step_info {
  int countP;        /* Poisson iterations (inner iterations in total when mixed) */
  double residualP;  /* relative residual at exit */
  int outer;         /* defect corrections (mixed) */
  int status;        /* 0 ok, 1 NaN/Inf residual */
  int countU[3];     /* implicit U: BiCGSTAB iterations per component */
  double residualU[3];
} ffvhc_step_info;

/* One Solver::Update step with explicit U (UpdateUXe/UYe/UZe 3535-3744, UpdateP 5294-5688,
 * UpdateU 5690-5771) for every selected case; each case advances its own step counter
 * (step 0 primes the Adams-Bashforth history).  info[c] per selected case; stage_ms (7 doubles
 * or NULL) = device ms of {total, -, UX, UY+UZ (0: fused into UX), -, P, U} as Solver::times.
 * Temperature (stage T) is out of scope. */
int ffvhc_step(ffvhc_ctx* ctx, ffvhc_step_fields* fields, const ffvhc_step_params* params, ffvhc_step_info* info,
               double* stage_ms);
/* Pressure force on the faces with boundary id cut_id (bcut_calc_f_p, bcut.f90:4666-4838, as summed
 * by Solver::PrintForceCID, Solver.cpp:2994-3160): f[3*i .. 3*i+2] for the i-th selected case.
 * p's ghost layer must be current. */
int ffvhc_force_p(ffvhc_ctx* ctx, int p, int cut_id, double* f);
/* Viscous force on the faces with boundary id cut_id (bcut_calc_f_v, bcut.f90:4840-5032, as summed
 * by Solver::PrintForceCID with the case's mu): f[3*i .. 3*i+2] for the i-th selected case.  The
 * ghost layers of ux, uy, uz must be current. */
int ffvhc_force_v(ffvhc_ctx* ctx, int ux, int uy, int uz, int cut_id, double* f);
/* No counterpart in the reference.  ffvhc_force_p for callers that cannot allocate device memory
 * (HIP graph capture, tight serving loops that read the drag every step, harnesses that forbid
 * hipMalloc): the same kernels and the same result, with the block partials and the per-case
 * triples in a fixed per-device scratch array instead of a per-call allocation.  It rejects a
 * context of more than FFVHC_NOALLOC_SLOTS block slots (status -1).  Every call on one GPU shares
 * that GPU's scratch, so calls that overlap on one GPU (two host threads, two contexts on the same
 * device) race; calls in sequence are independent (every value read is written by the same call). */
#define FFVHC_NOALLOC_SLOTS 131072
int ffvhc_force_p_noalloc(ffvhc_ctx* ctx, int p, int cut_id, double* f);

/* LocalScalar3D::CalcStats (LocalScalar3D.cpp:9-80): out[5*i ..] = {sum, max, min, |max|, |min|}
 * over the interior cells of the i-th selected case. */
int ffvhc_field_stats(ffvhc_ctx* ctx, int field, double* out);
/* The CFL guard of Solver::PrintStats (Solver.cpp:2655-2676): cfl[i] = |ux|max dt / dx with
 * dx = RootBlockLength / 2^maxLevel; returns 2 when any selected case exceeds 0.9 (the reference
 * exits there). */
int ffvhc_cfl_check(ffvhc_ctx* ctx, int ux, double* cfl);

/* Step counter and inflow ramp (TimeControl/Acceleration, Solver.cpp:3347-3355) of a case. */
int ffvhc_case_set_step(ffvhc_ctx* ctx, int case_id, int step, int accel_steps);
int ffvhc_case_step(const ffvhc_ctx* ctx, int case_id);

#ifdef __cplusplus
}
#endif
#endif
