// Device-side data model shared by every kernel.
//
// HBM layout (one MI355X holds every block of every octree level of every case):
//   field interior : T[slot][16][16][16]    (k, j, i; i fastest: one 16-double row = one 128 B line)
//   field halo     : T[slot][6][16*16]      (face X_M..Z_P; the one ghost layer every 7-point
//                                            stencil reads; position a + 16*b over the face's
//                                            tangential axes (j,k) for X, (i,k) for Y, (i,j) for Z)
// The reference's padded 20^3 per-block arrays (vc = 2, LocalScalar3D.h:232-235) are the host
// ABI format only; the second ghost layer is never read by any 7-point stencil and is not stored.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace ffvhc {

constexpr int NB = 16;               // cells per block edge
constexpr int NC = NB * NB * NB;     // interior cells per block
constexpr int NF = NB * NB;          // cells per face
constexpr int NH = 6 * NF;           // halo cells per block
constexpr int MAX_CASES = 1024;
constexpr int MAX_FIELDS = 512;

// Outer-BC groups carried by each case: the velocity components, pressure and the
// boundaryTypeNULL fields (Neumann 0: A, b, face velocities, scratch, Solver.cpp:2024-2034).
// BC_HOM carries the pressure BC types with zero values: the homogeneous BC of a correction.
enum BcGroup { BC_UX = 0, BC_UY = 1, BC_UZ = 2, BC_P = 3, BC_NULL = 4, BC_ZERO = 5, BC_HOM = 6, NUM_BC = 7 };
constexpr int NUM_USER_BC = 6;
enum BcType { BCT_NONE = -1, BCT_DIRICHLET = 0, BCT_NEUMANN = 1, BCT_PERIODIC = 9 };

struct DevBlock {
  int case_id;
  int level;
  double dx;          // cell size
  double org[3];      // block origin
  int ftype[6];       // FaceType
  int nbr[6][4];      // neighbour slots
  int sub[6];         // this block's quarter on a coarser neighbour's face
  int flags;          // bit 0: block holds cut or solid cells
  int pad;
};

struct DevCase {
  int slot_begin;
  int nblk;
  double rho, mu, dt;
  double gx, gy, gz, betag, tr;
  int bct[NUM_BC][6];
  double bcv[NUM_BC][6];
  int step;  // time step counter of the case (0 primes the Adams-Bashforth history)
  int pad_;
};

// Device pointers of the time-step fields (FP64, interior + one ghost layer where a stencil
// reads neighbours).  uc/ud receive this step's convection/diffusion terms, ucp/udp hold the
// previous step's (the caller swaps the roles after the step instead of copying).
struct StepFields {
  double* u[3];
  double* hu[3];
  double* p;
  double* hp;
  double* uc[3];
  double* ud[3];
  const double* ucp[3];
  const double* udp[3];
  double* v[6];
  double* A[7];
  double* b;
  double* lapp;
  const double* t;     // temperature (buoyancy) or null = 0
  const double* f[3];  // body force or null = 0
};

// Per-case state of an iterative solve, updated on the device by the per-case finalisation of
// each reduction (deterministic: partials are summed in slot order).
struct SolveState {
  double bb;        // b.b
  double rr;        // last residual norm squared
  double rr0, rr1;  // CG
  double alpha, beta, gamma;
  double residual;  // FFVILS IsConverged residual
  double omega, eps;
  int countMax;
  int count;        // reference `count` semantics
  int done;         // 1 once converged / exhausted / failed
  int status;       // 0 ok, 1 NaN/Inf residual (reference Exit(0))
  int parity;       // buffer that holds the result (RBGS ping-pong)
  int iter;         // iterations launched so far
  int restart;      // Krylov restart pending: the next P step takes p = r (reliable update)
  int used;         // outer state: inner iterations spent so far (budget of a restarted inner solve)
};

// the seven bands of a 7-point matrix {Ap, Aw, Ae, As, An, Ab, At}
template <class T>
struct Coef7 {
  const T* p[7];
};

// BiCGSTAB work vectors r r0 p p_ q_ s s_ t_ and the ghost layers of p_ and s_
template <class T>
struct BiVecs {
  T *r, *r0, *p, *pp, *q, *s, *ss, *t;
  T *hpp, *hss;
};

__host__ __device__ inline int halo_index(int face, int a, int b) { return face * NF + a + NB * b; }

// cell (i,j,k) that touches face `face` at face position (a,b), layer `l` (0 = adjacent)
__host__ __device__ inline void face_cell(int face, int a, int b, int l, int& i, int& j, int& k) {
  const bool plus = face & 1;
  const int nrm = plus ? NB - 1 - l : l;
  switch (face >> 1) {
    case 0: i = nrm; j = a; k = b; break;
    case 1: i = a; j = nrm; k = b; break;
    default: i = a; j = b; k = nrm; break;
  }
}

// Face-position axes: the coordinates (a,b) of a face map to (t0,t1) axes used by subface
// numbering (X: y,z  Y: z,x  Z: x,y); a/b themselves are (j,k), (i,k), (i,j).
__host__ __device__ inline void face_ab_to_o(int face, int a, int b, int& u0, int& u1) {
  // returns the coordinates along subface axes origin0 / origin1
  switch (face >> 1) {
    case 0: u0 = a; u1 = b; break;  // (y, z)
    case 1: u0 = b; u1 = a; break;  // (z, x)
    default: u0 = a; u1 = b; break; // (x, y)
  }
}

}  // namespace ffvhc
