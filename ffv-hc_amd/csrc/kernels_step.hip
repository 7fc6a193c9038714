// Stage kernels of the fractional step (Solver::Update, Solver.cpp:3344-3416) with explicit U:
//
//   predict  : per velocity component d, bcut_calc_c_f_c2 (convection, bcut.f90:1583-1764),
//              bcut_calc_d_u (diffusion, 3397-3552) and bcut_update_u (Adams-Bashforth 2,
//              4299-4523) fused; step 0 primes ucp/udp with this step's terms (Solver.cpp:3693-3702)
//   p1       : bcut_add_g (3793-3833), bcut_add_f (3835-3877), bcut_remove_p (3879-4061) fused
//   ab_p     : bcut_calc_ab_p (608-832): Poisson matrix, b = div(v)/dt, face velocities
//   refval   : bcut_set_reference_value (5626-5710), including its Ae(i+1) index at 5682
//   matrix BC: bc_a*_{d,n} (bca.f90), face by face in face order (BlockScalar3D.h:183-190)
//   corr     : bcut_corr_u (4063-4297): projection, face-velocity correction, lapp
//
// One 256-thread workgroup per 16^3 block, thread = (i, j) column over the 16 k planes. Every
// cell evaluates the reference's expressions in the reference's order (FP64), so results are
// bit-identical to the per-block Fortran.  Geometry per cell: six float32 cut distances, a
// flag word (bit f: face f is cut, bit 6: fluid), the cut ids only for forces.
#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

constexpr int ST_NT = 256;

// value of a field at (i, j, k) with at most one coordinate one cell outside the block
__device__ inline double at(const double* __restrict__ x, const double* __restrict__ h, int slot, int i, int j,
                            int k) {
  const size_t hb = (size_t)slot * NH;
  if (i < 0) return h[hb + halo_index(0, j, k)];
  if (i >= NB) return h[hb + halo_index(1, j, k)];
  if (j < 0) return h[hb + halo_index(2, i, k)];
  if (j >= NB) return h[hb + halo_index(3, i, k)];
  if (k < 0) return h[hb + halo_index(4, i, j)];
  if (k >= NB) return h[hb + halo_index(5, i, j)];
  return x[(size_t)slot * NC + i + NB * (j + NB * k)];
}

struct Geo {
  double d[6];
  uint32_t fl;
  __device__ bool cut(int f) const { return (fl >> f) & 1u; }
  __device__ bool fluid() const { return (fl >> 6) & 1u; }
};

__device__ inline Geo load_geo(const float* __restrict__ cut, const uint32_t* __restrict__ flag, size_t capnc,
                               size_t g) {
  Geo G;
#pragma unroll
  for (int f = 0; f < 6; f++) G.d[f] = (double)cut[(size_t)f * capnc + g];
  G.fl = flag[g];
  return G;
}

struct Face6 {
  double x0, x1, y2, y3, z4, z5;
};

// pressure-gradient face terms shared by update_u / remove_p / corr_u (bcut.f90), with the
// "- g*0.0" terms of update_u when with_g
__device__ inline Face6 grad_faces(const double* __restrict__ p, const double* __restrict__ hp, int slot, int i,
                                   int j, int k, double r0, double dx, bool with_g, double gx, double gy, double gz,
                                   const Geo& G) {
  const double pc = at(p, hp, slot, i, j, k);
  const double dpx0 = (pc - at(p, hp, slot, i - 1, j, k)) / dx, dpx1 = (at(p, hp, slot, i + 1, j, k) - pc) / dx;
  const double dpy2 = (pc - at(p, hp, slot, i, j - 1, k)) / dx, dpy3 = (at(p, hp, slot, i, j + 1, k) - pc) / dx;
  const double dpz4 = (pc - at(p, hp, slot, i, j, k - 1)) / dx, dpz5 = (at(p, hp, slot, i, j, k + 1) - pc) / dx;
  Face6 r;
  if (with_g) {
    r.x0 = r0 * dpx0 - gx * 0.0;
    r.x1 = r0 * dpx1 - gx * 0.0;
    r.y2 = r0 * dpy2 - gy * 0.0;
    r.y3 = r0 * dpy3 - gy * 0.0;
    r.z4 = r0 * dpz4 - gz * 0.0;
    r.z5 = r0 * dpz5 - gz * 0.0;
  } else {
    r.x0 = r0 * dpx0;
    r.x1 = r0 * dpx1;
    r.y2 = r0 * dpy2;
    r.y3 = r0 * dpy3;
    r.z4 = r0 * dpz4;
    r.z5 = r0 * dpz5;
  }
  const double* d = G.d;
  if (G.cut(0)) r.x0 = r.x1 * (d[0] - 0.5) / (d[0] + 0.5);
  if (G.cut(1)) r.x1 = r.x0 * (d[1] - 0.5) / (d[1] + 0.5);
  if (G.cut(2)) r.y2 = r.y3 * (d[2] - 0.5) / (d[2] + 0.5);
  if (G.cut(3)) r.y3 = r.y2 * (d[3] - 0.5) / (d[3] + 0.5);
  if (G.cut(4)) r.z4 = r.z5 * (d[4] - 0.5) / (d[4] + 0.5);
  if (G.cut(5)) r.z5 = r.z4 * (d[5] - 0.5) / (d[5] + 0.5);
  if (G.cut(0) && G.cut(1)) { r.x0 = 0.0; r.x1 = 0.0; }
  if (G.cut(2) && G.cut(3)) { r.y2 = 0.0; r.y3 = 0.0; }
  if (G.cut(4) && G.cut(5)) { r.z4 = 0.0; r.z5 = 0.0; }
  return r;
}

// neighbour values of u around a cell: w e s n b t
struct Nb6 {
  double p, w, e, s, n, b, t;
};

__device__ inline Nb6 load_nb(const double* __restrict__ u, const double* __restrict__ hu, int slot, int i, int j,
                              int k) {
  Nb6 v;
  v.p = u[(size_t)slot * NC + i + NB * (j + NB * k)];
  v.w = at(u, hu, slot, i - 1, j, k);
  v.e = at(u, hu, slot, i + 1, j, k);
  v.s = at(u, hu, slot, i, j - 1, k);
  v.n = at(u, hu, slot, i, j + 1, k);
  v.b = at(u, hu, slot, i, j, k - 1);
  v.t = at(u, hu, slot, i, j, k + 1);
  return v;
}

// bcut_calc_c_f_c2 (bcut.f90:1583-1764) at one cell
__device__ inline double conv_c2(const Nb6& f, const double* const* V, size_t g, const Geo& G, double dx) {
  const double* d = G.d;
  const double vx0 = V[0][g], vx1 = V[1][g], vy2 = V[2][g], vy3 = V[3][g], vz4 = V[4][g], vz5 = V[5][g];
  const double f0 = 0.5 * (f.p + f.w), f1 = 0.5 * (f.p + f.e), f2 = 0.5 * (f.p + f.s);
  const double f3 = 0.5 * (f.p + f.n), f4 = 0.5 * (f.p + f.b), f5 = 0.5 * (f.p + f.t);
  double q0 = vx0 * f0, q1 = vx1 * f1, q2 = vy2 * f2, q3 = vy3 * f3, q4 = vz4 * f4, q5 = vz5 * f5;
  if (G.cut(0)) q0 = (d[0] - 0.5) / (d[0] + 0.5) * q1;
  if (G.cut(1)) q1 = (d[1] - 0.5) / (d[1] + 0.5) * q0;
  if (G.cut(2)) q2 = (d[2] - 0.5) / (d[2] + 0.5) * q3;
  if (G.cut(3)) q3 = (d[3] - 0.5) / (d[3] + 0.5) * q2;
  if (G.cut(4)) q4 = (d[4] - 0.5) / (d[4] + 0.5) * q5;
  if (G.cut(5)) q5 = (d[5] - 0.5) / (d[5] + 0.5) * q4;
  if (G.cut(0) && G.cut(1)) { q0 = 0.0; q1 = 0.0; }
  if (G.cut(2) && G.cut(3)) { q2 = 0.0; q3 = 0.0; }
  if (G.cut(4) && G.cut(5)) { q4 = 0.0; q5 = 0.0; }
  double fc = (q1 - q0) / dx + (q3 - q2) / dx + (q5 - q4) / dx;
  if (!G.fluid()) fc = 0.0;
  return fc;
}

// cut-face viscosities and wall masks shared by calc_d_u and calc_abd_u (bcut.f90)
struct Visc {
  double mu[6], m[6];
};

__device__ inline Visc cut_visc(double mu, const Geo& G) {
  const double* d = G.d;
  Visc v;
  for (int q = 0; q < 6; q++) {
    v.mu[q] = mu;
    v.m[q] = 0.0;
  }
  if (G.cut(0)) { v.mu[0] = mu / d[0] * 2.0 / (d[0] + d[1]); v.mu[1] = mu / d[1] * 2.0 / (d[0] + d[1]); v.m[0] = 1.0; }
  if (G.cut(1)) { v.mu[0] = mu / d[0] * 2.0 / (d[0] + d[1]); v.mu[1] = mu / d[1] * 2.0 / (d[0] + d[1]); v.m[1] = 1.0; }
  if (G.cut(2)) { v.mu[2] = mu / d[2] * 2.0 / (d[2] + d[3]); v.mu[3] = mu / d[3] * 2.0 / (d[2] + d[3]); v.m[2] = 1.0; }
  if (G.cut(3)) { v.mu[2] = mu / d[2] * 2.0 / (d[2] + d[3]); v.mu[3] = mu / d[3] * 2.0 / (d[2] + d[3]); v.m[3] = 1.0; }
  if (G.cut(4)) { v.mu[4] = mu / d[4] * 2.0 / (d[4] + d[5]); v.mu[5] = mu / d[5] * 2.0 / (d[4] + d[5]); v.m[4] = 1.0; }
  if (G.cut(5)) { v.mu[4] = mu / d[4] * 2.0 / (d[4] + d[5]); v.mu[5] = mu / d[5] * 2.0 / (d[4] + d[5]); v.m[5] = 1.0; }
  return v;
}

// diffusion sum l1 (ue - up) - l0 (up - uw) + ... with wall value Uc behind masked faces
__device__ inline double diff_sum(const Nb6& f, const double l[6], const double m[6], double Uc) {
  const double uw = f.w * (1.0 - m[0]) + Uc * m[0];
  const double ue = f.e * (1.0 - m[1]) + Uc * m[1];
  const double us = f.s * (1.0 - m[2]) + Uc * m[2];
  const double un = f.n * (1.0 - m[3]) + Uc * m[3];
  const double ub = f.b * (1.0 - m[4]) + Uc * m[4];
  const double ut = f.t * (1.0 - m[5]) + Uc * m[5];
  return l[1] * (ue - f.p) - l[0] * (f.p - uw) + l[3] * (un - f.p) - l[2] * (f.p - us) + l[5] * (ut - f.p) -
         l[4] * (f.p - ub);
}

// ------------------------------------------------------------------ U predictor (explicit)
__global__ __launch_bounds__(ST_NT) void predict_kernel(const int* __restrict__ active,
                                                        const DevBlock* __restrict__ blk,
                                                        const DevCase* __restrict__ cs, StepFields F,
                                                        const float* __restrict__ cut,
                                                        const uint32_t* __restrict__ flag, size_t capnc) {
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx, dt = C.dt, rhof = C.rho, mu = C.mu, Uc = 0.0;
  const bool step0 = C.step == 0;
  {
    // one workgroup per block and component (the components are independent: each reads its own
    // u, the face velocities v and p, and writes its own uc, ud and u), which triples the
    // workgroups of small cases
    const int dir = blockIdx.y;
    const double* __restrict__ u = F.u[dir];
    const double* __restrict__ hu = F.hu[dir];
    double unew[NB];
    for (int k = 0; k < NB; k++) {
      const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
      const Geo G = load_geo(cut, flag, capnc, g);
      const bool fluid = G.fluid();
      const Nb6 f = load_nb(u, hu, slot, i, j, k);
      const double fp = f.p;
      const double fc = conv_c2(f, F.v, g, G, dx);  // bcut_calc_c_f_c2
      double ud;                                     // bcut_calc_d_u
      {
        const Visc vs = cut_visc(mu, G);
        double l[6];
        for (int q = 0; q < 6; q++) l[q] = vs.mu[q] / rhof / (dx * dx);
        ud = diff_sum(f, l, vs.m, Uc);
        if (!fluid) ud = 0.0;
      }
      // bcut_update_u
      const double ucp = step0 ? fc : F.ucp[dir][g];
      const double udp = step0 ? ud : F.udp[dir][g];
      const Face6 r = grad_faces(F.p, F.hp, slot, i, j, k, 1.0 / rhof, dx, true, C.gx, C.gy, C.gz, G);
      const double rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
      const double rdp = dir == 0 ? rdpx : (dir == 1 ? rdpy : rdpz);
      double un = fp - (1.5 * fc - 0.5 * ucp) * dt + (1.5 * ud - 0.5 * udp) * dt - rdp * dt;
      if (!fluid) un = Uc;
      unew[k] = un;
      F.uc[dir][g] = fc;
      F.ud[dir][g] = ud;
    }
    __syncthreads();  // every neighbour read of u[dir] in this block is done
    for (int k = 0; k < NB; k++) F.u[dir][(size_t)slot * NC + i + NB * (j + NB * k)] = unew[k];
  }
}

// ------------------------------------------------------------------ U implicit: matrix and RHS
// UpdateUX/UY/UZ (Solver.cpp:4391-4690) for component dir: bcut_calc_c_f_c2 into uc,
// bcut_calc_d_u, then bcut_calc_abd_u (bcut.f90:43-322: A = 1 + alpha sum(l), b with the
// Adams-Bashforth convection, (1 - alpha) diffusion and the pressure gradient; ud overwritten;
// solid cells get the identity row, b = Uc and u = 0).  ucp is this step's previous convection
// (zero before the first step, as the initial Fill leaves it).
__global__ __launch_bounds__(ST_NT) void abd_u_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                      const DevCase* __restrict__ cs, StepFields F,
                                                      const float* __restrict__ cut, const uint32_t* __restrict__ flag,
                                                      size_t capnc, int dir, double alpha) {
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx, dt = C.dt, rhof = C.rho, mu = C.mu, Uc = 0.0;
  const bool step0 = C.step == 0;
  const double* __restrict__ u = F.u[dir];
  const double* __restrict__ hu = F.hu[dir];
  double unew[NB];
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    const Geo G = load_geo(cut, flag, capnc, g);
    const bool fluid = G.fluid();
    const Nb6 f = load_nb(u, hu, slot, i, j, k);
    const double fc = conv_c2(f, F.v, g, G, dx);
    // (bcut_calc_d_u's result is overwritten by calc_abd_u below)
    Visc vs = cut_visc(mu, G);
    const Face6 r = grad_faces(F.p, F.hp, slot, i, j, k, 1.0 / rhof, dx, true, C.gx, C.gy, C.gz, G);
    if (G.cut(0) && G.cut(1)) { vs.m[0] = 1.0; vs.m[1] = 1.0; }
    if (G.cut(2) && G.cut(3)) { vs.m[2] = 1.0; vs.m[3] = 1.0; }
    if (G.cut(4) && G.cut(5)) { vs.m[4] = 1.0; vs.m[5] = 1.0; }
    const double rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
    const double rdp = dir == 0 ? rdpx : (dir == 1 ? rdpy : rdpz);
    double l[6];
    for (int q = 0; q < 6; q++) l[q] = vs.mu[q] / rhof * dt / (dx * dx);
    // the reference loop zeroes u of solid cells as it goes (u(i,j,k) = 0 at the end of each
    // cell), so the w, s and b neighbours inside the block, already visited, read 0 if solid
    Nb6 fz = f;
    const size_t g0 = (size_t)slot * NC;
    if (i > 0 && !((flag[g0 + (i - 1) + NB * (j + NB * k)] >> 6) & 1u)) fz.w = 0.0;
    if (j > 0 && !((flag[g0 + i + NB * ((j - 1) + NB * k)] >> 6) & 1u)) fz.s = 0.0;
    if (k > 0 && !((flag[g0 + i + NB * (j + NB * (k - 1))] >> 6) & 1u)) fz.b = 0.0;
    double ud = diff_sum(fz, l, vs.m, Uc);
    if (!fluid) ud = 0.0;
    const double* m = vs.m;
    double a[7];
    a[0] = 1.0 + alpha * (l[0] + l[1]) + alpha * (l[2] + l[3]) + alpha * (l[4] + l[5]);
    for (int q = 0; q < 6; q++) a[1 + q] = -(alpha * l[q] * (1.0 - m[q]));
    const double ucp = step0 ? 0.0 : F.ucp[dir][g];
    double bv = f.p - (1.5 * fc - 0.5 * ucp) * dt + (1.0 - alpha) * ud * dt - rdp * dt + alpha * l[0] * Uc * m[0] +
                alpha * l[1] * Uc * m[1] + alpha * l[2] * Uc * m[2] + alpha * l[3] * Uc * m[3] +
                alpha * l[4] * Uc * m[4] + alpha * l[5] * Uc * m[5];
    double un = f.p;
    if (!fluid) {
      a[0] = 1.0;
      for (int q = 1; q < 7; q++) a[q] = 0.0;
      bv = Uc;
      un = 0.0;
    }
    for (int q = 0; q < 7; q++) F.A[q][g] = a[q];
    F.b[g] = bv;
    F.uc[dir][g] = fc;
    F.ud[dir][g] = ud;
    unew[k] = un;
  }
  __syncthreads();  // every neighbour read of u[dir] in this block is done
  for (int k = 0; k < NB; k++) F.u[dir][(size_t)slot * NC + i + NB * (j + NB * k)] = unew[k];
}

// ------------------------------------------------------------------ add_g, add_f, remove_p
__global__ __launch_bounds__(ST_NT) void p1_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                   const DevCase* __restrict__ cs, StepFields F,
                                                   const float* __restrict__ cut, const uint32_t* __restrict__ flag,
                                                   size_t capnc) {
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx, dt = C.dt, rhof = C.rho;
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    const Geo G = load_geo(cut, flag, capnc, g);
    double ux = F.u[0][g], uy = F.u[1][g], uz = F.u[2][g];
    const double tv = F.t ? F.t[g] : 0.0;
    // bcut_add_g
    ux = ux + C.gx * dt;
    uy = uy + C.gy * dt;
    uz = uz + C.gz * dt + C.betag * (tv - C.tr) * dt;
    // bcut_add_f (fluid cells)
    if (G.fluid()) {
      const double fx = F.f[0] ? F.f[0][g] : 0.0, fy = F.f[1] ? F.f[1][g] : 0.0, fz = F.f[2] ? F.f[2][g] : 0.0;
      ux = ux + fx / rhof * dt;
      uy = uy + fy / rhof * dt;
      uz = uz + fz / rhof * dt;
    }
    // bcut_remove_p
    const Face6 r = grad_faces(F.p, F.hp, slot, i, j, k, 1.0 / rhof, dx, false, 0.0, 0.0, 0.0, G);
    const double rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
    ux = ux + rdpx * dt;
    uy = uy + rdpy * dt;
    uz = uz + rdpz * dt;
    if (!G.fluid()) { ux = 0.0; uy = 0.0; uz = 0.0; }
    F.u[0][g] = ux;
    F.u[1][g] = uy;
    F.u[2][g] = uz;
  }
}

// ------------------------------------------------------------------ Poisson matrix and RHS
__global__ __launch_bounds__(ST_NT) void ab_p_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                     const DevCase* __restrict__ cs, StepFields F,
                                                     const float* __restrict__ cut, const uint32_t* __restrict__ flag,
                                                     size_t capnc) {
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx, dt = C.dt, rhof = C.rho;
  const double *ux = F.u[0], *uy = F.u[1], *uz = F.u[2];
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    const Geo G = load_geo(cut, flag, capnc, g);
    const double* d = G.d;
    double r0 = 1.0 / rhof, r1 = 1.0 / rhof, r2 = 1.0 / rhof, r3 = 1.0 / rhof, r4 = 1.0 / rhof, r5 = 1.0 / rhof;
    double vx0 = 0.5 * (ux[g] + at(ux, F.hu[0], slot, i - 1, j, k));
    double vx1 = 0.5 * (ux[g] + at(ux, F.hu[0], slot, i + 1, j, k));
    double vy2 = 0.5 * (uy[g] + at(uy, F.hu[1], slot, i, j - 1, k));
    double vy3 = 0.5 * (uy[g] + at(uy, F.hu[1], slot, i, j + 1, k));
    double vz4 = 0.5 * (uz[g] + at(uz, F.hu[2], slot, i, j, k - 1));
    double vz5 = 0.5 * (uz[g] + at(uz, F.hu[2], slot, i, j, k + 1));
    if (G.cut(0)) { r0 = 0.0; vx0 = vx1 * (d[0] - 0.5) / (d[0] + 0.5); r1 = (1.0 / rhof) / (d[0] + 0.5); }
    if (G.cut(1)) { r1 = 0.0; vx1 = vx0 * (d[1] - 0.5) / (d[1] + 0.5); r0 = (1.0 / rhof) / (d[1] + 0.5); }
    if (G.cut(2)) { r2 = 0.0; vy2 = vy3 * (d[2] - 0.5) / (d[2] + 0.5); r3 = (1.0 / rhof) / (d[2] + 0.5); }
    if (G.cut(3)) { r3 = 0.0; vy3 = vy2 * (d[3] - 0.5) / (d[3] + 0.5); r2 = (1.0 / rhof) / (d[3] + 0.5); }
    if (G.cut(4)) { r4 = 0.0; vz4 = vz5 * (d[4] - 0.5) / (d[4] + 0.5); r5 = (1.0 / rhof) / (d[4] + 0.5); }
    if (G.cut(5)) { r5 = 0.0; vz5 = vz4 * (d[5] - 0.5) / (d[5] + 0.5); r4 = (1.0 / rhof) / (d[5] + 0.5); }
    if (G.cut(0) && G.cut(1)) { r0 = 0.0; r1 = 0.0; vx0 = 0.0; vx1 = 0.0; }
    if (G.cut(2) && G.cut(3)) { r2 = 0.0; r3 = 0.0; vy2 = 0.0; vy3 = 0.0; }
    if (G.cut(4) && G.cut(5)) { r4 = 0.0; r5 = 0.0; vz4 = 0.0; vz5 = 0.0; }
    const double l0 = r0 / (dx * dx), l1 = r1 / (dx * dx), l2 = r2 / (dx * dx);
    const double l3 = r3 / (dx * dx), l4 = r4 / (dx * dx), l5 = r5 / (dx * dx);
    const double divv = (vx1 - vx0 + vy3 - vy2 + vz5 - vz4) / dx;
    double a[7] = {-(l0 + l1) - (l2 + l3) - (l4 + l5), l0, l1, l2, l3, l4, l5};
    double bv = divv / dt;
    double v[6] = {vx0, vx1, vy2, vy3, vz4, vz5};
    if (!G.fluid()) {
      a[0] = -6.0;
      for (int q = 1; q < 7; q++) a[q] = 1.0;
      bv = 0.0;
      for (int q = 0; q < 6; q++) v[q] = 0.0;
    }
    for (int q = 0; q < 7; q++) F.A[q][g] = a[q];
    F.b[g] = bv;
    for (int q = 0; q < 6; q++) F.v[q][g] = v[q];
  }
}

// ------------------------------------------------------------------ reference pressure
// One thread per candidate block walks the block's cells in the reference loop order, so a
// point on a cell boundary is applied to each matching cell in turn exactly as the reference
// does; writes that land in the ghost layer are dropped (no interior loop reads them).
__global__ void refval_kernel(const int* __restrict__ slots, const DevBlock* __restrict__ blk, StepFields F,
                              double xr, double yr, double zr, double pr) {
  if (threadIdx.x != 0) return;
  const int slot = slots[blockIdx.x];
  const DevBlock& B = blk[slot];
  const double dx = B.dx;
  double* A[7];
  for (int q = 0; q < 7; q++) A[q] = F.A[q] + (size_t)slot * NC;
  double* b = F.b + (size_t)slot * NC;
  auto in = [](int i, int j, int k) { return i >= 0 && i < NB && j >= 0 && j < NB && k >= 0 && k < NB; };
  auto ix = [](int i, int j, int k) { return i + NB * (j + NB * k); };
  for (int k = 0; k < NB; k++)
    for (int j = 0; j < NB; j++)
      for (int i = 0; i < NB; i++) {
        // Fortran cell (i+1, j+1, k+1): [org + (i) dx, org + (i+1) dx]
        const double x0 = B.org[0] + (double)i * dx, y0 = B.org[1] + (double)j * dx, z0 = B.org[2] + (double)k * dx;
        const double x1 = B.org[0] + (double)(i + 1) * dx, y1 = B.org[1] + (double)(j + 1) * dx,
                     z1 = B.org[2] + (double)(k + 1) * dx;
        if (!((x0 <= xr && xr <= x1) && (y0 <= yr && yr <= y1) && (z0 <= zr && zr <= z1))) continue;
        const int m = ix(i, j, k);
        A[0][m] = 1.0;
        for (int q = 1; q < 7; q++) A[q][m] = 0.0;
        b[m] = pr;
        // neighbours e, w, n, s, t, b; the w line reads Ae of the e cell (bcut.f90:5682)
        const int ne[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
        const int band[6] = {1, 2, 3, 4, 5, 6};  // Aw of e, Ae of w, As of n, An of s, Ab of t, At of b
        for (int q = 0; q < 6; q++) {
          const int a0 = ne[q][0], a1 = ne[q][1], a2 = ne[q][2];
          if (!in(a0, a1, a2)) continue;
          const int n = ix(a0, a1, a2);
          // the w line subtracts Ae of the e cell; when e lies in the ghost layer that matrix
          // entry is never assigned (calc_ab_p writes interior cells only) and stays zero
          const double ap_sub = q == 1 ? (in(i + 1, j, k) ? A[2][ix(i + 1, j, k)] : 0.0) : A[band[q]][n];
          A[0][n] = A[0][n] - ap_sub;
          b[n] = b[n] - A[band[q]][n] * pr * 2.0;
          A[band[q]][n] = 0.0;
        }
      }
}

// ------------------------------------------------------------------ matrix outer BC
__global__ __launch_bounds__(ST_NT) void matrix_bc_kernel(const int* __restrict__ active,
                                                          const DevBlock* __restrict__ blk,
                                                          const DevCase* __restrict__ cs, StepFields F, int bcg) {
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int t = threadIdx.x;
  for (int f = 0; f < 6; f++) {
    if (B.ftype[f] != FACE_OUTER) continue;
    const int type = C.bct[bcg][f];
    if (type != BCT_DIRICHLET && type != BCT_NEUMANN) continue;
    const double xc = type == BCT_NEUMANN ? C.bcv[bcg][f] * B.dx : C.bcv[bcg][f];
    int i, j, k;
    face_cell(f, t & 15, t >> 4, 0, i, j, k);
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    double* Ap = F.A[0];
    double* An = F.A[1 + f];
    if (type == BCT_DIRICHLET) {
      Ap[g] = Ap[g] - An[g];
      F.b[g] = F.b[g] - An[g] * xc * 2.0;
    } else {
      Ap[g] = Ap[g] + An[g];
      F.b[g] = (f & 1) ? F.b[g] - An[g] * xc : F.b[g] + An[g] * xc;
    }
    An[g] = 0.0;
    __syncthreads();  // faces in order: a corner cell folds face f before face f+1
  }
}

// ------------------------------------------------------------------ velocity correction
__global__ __launch_bounds__(ST_NT) void corr_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                     const DevCase* __restrict__ cs, StepFields F,
                                                     const float* __restrict__ cut, const uint32_t* __restrict__ flag,
                                                     size_t capnc) {
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const DevCase& C = cs[B.case_id];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx, dt = C.dt, rhof = C.rho;
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    const Geo G = load_geo(cut, flag, capnc, g);
    const Face6 r = grad_faces(F.p, F.hp, slot, i, j, k, 1.0 / rhof, dx, false, 0.0, 0.0, 0.0, G);
    const double rdpx = 0.5 * (r.x0 + r.x1), rdpy = 0.5 * (r.y2 + r.y3), rdpz = 0.5 * (r.z4 + r.z5);
    double ux = F.u[0][g] - rdpx * dt, uy = F.u[1][g] - rdpy * dt, uz = F.u[2][g] - rdpz * dt;
    double v[6] = {F.v[0][g] - r.x0 * dt, F.v[1][g] - r.x1 * dt, F.v[2][g] - r.y2 * dt,
                   F.v[3][g] - r.y3 * dt, F.v[4][g] - r.z4 * dt, F.v[5][g] - r.z5 * dt};
    const double pc = F.p[g];
    double lapp = (at(F.p, F.hp, slot, i + 1, j, k) + at(F.p, F.hp, slot, i - 1, j, k) +
                   at(F.p, F.hp, slot, i, j + 1, k) + at(F.p, F.hp, slot, i, j - 1, k) +
                   at(F.p, F.hp, slot, i, j, k + 1) + at(F.p, F.hp, slot, i, j, k - 1) - 6.0 * pc) /
                  (dx * dx);
    if (!G.fluid()) {
      ux = 0.0;
      uy = 0.0;
      uz = 0.0;
      for (int q = 0; q < 6; q++) v[q] = 0.0;
      lapp = 0.0;
    }
    if (G.fl & 63u) lapp = 0.0;
    F.u[0][g] = ux;
    F.u[1][g] = uy;
    F.u[2][g] = uz;
    for (int q = 0; q < 6; q++) F.v[q][g] = v[q];
    F.lapp[g] = lapp;
  }
}

// ------------------------------------------------------------------ pressure force
// bcut_calc_f_p (bcut.f90:4666-4838) summed as Solver::PrintForceCID does (Solver.cpp:2994-3160):
// on every face of a fluid cell whose boundary id is cid_target the face pressure is extrapolated
// from the cell and its opposite neighbour, force += (m1 p1 - m0 p0, ...) dx^2.  One partial
// triple per block, then fixed-order per-case sums in force_sum_kernel.
__global__ __launch_bounds__(ST_NT) void force_p_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                        const double* __restrict__ p, const double* __restrict__ hp,
                                                        const uint8_t* __restrict__ cid8,
                                                        const uint32_t* __restrict__ flag, size_t capnc, int target,
                                                        double* __restrict__ part, size_t nslots) {
  __shared__ double red[ST_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    const double pp = p[g];
    const double pw = at(p, hp, slot, i - 1, j, k), pe = at(p, hp, slot, i + 1, j, k);
    const double ps = at(p, hp, slot, i, j - 1, k), pn = at(p, hp, slot, i, j + 1, k);
    const double pb = at(p, hp, slot, i, j, k - 1), pt = at(p, hp, slot, i, j, k + 1);
    double p0 = 0.5 * (pp + pw), p1 = 0.5 * (pp + pe), p2 = 0.5 * (pp + ps);
    double p3 = 0.5 * (pp + pn), p4 = 0.5 * (pp + pb), p5 = 0.5 * (pp + pt);
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0, m4 = 0.0, m5 = 0.0;
    if (cid8[0 * capnc + g] == target) { p0 = pp + (pp - pe) * 0.5; m0 = 1.0; }
    if (cid8[1 * capnc + g] == target) { p1 = pp + (pp - pw) * 0.5; m1 = 1.0; }
    if (cid8[2 * capnc + g] == target) { p2 = pp + (pp - pn) * 0.5; m2 = 1.0; }
    if (cid8[3 * capnc + g] == target) { p3 = pp + (pp - ps) * 0.5; m3 = 1.0; }
    if (cid8[4 * capnc + g] == target) { p4 = pp + (pp - pt) * 0.5; m4 = 1.0; }
    if (cid8[5 * capnc + g] == target) { p5 = pp + (pp - pb) * 0.5; m5 = 1.0; }
    if (!((flag[g] >> 6) & 1u)) { m0 = m1 = m2 = m3 = m4 = m5 = 0.0; }
    fx += (m1 * p1 - m0 * p0) * dx * dx;
    fy += (m3 * p3 - m2 * p2) * dx * dx;
    fz += (m5 * p5 - m4 * p4) * dx * dx;
  }
  const double sx = block_sum<ST_NT>(fx, red), sy = block_sum<ST_NT>(fy, red), sz = block_sum<ST_NT>(fz, red);
  if (t == 0) {
    part[slot] = sx;
    part[nslots + slot] = sy;
    part[2 * nslots + slot] = sz;
  }
}

// bcut_calc_f_v (bcut.f90:4840-5032): viscous force on the faces cut by `target`.  A cut face
// takes the wall-extrapolated neighbour (1 - 1/d) u_p; per component the doubled normal gradient
// on its own axis and, on the two others, the gradient plus half the cross difference of the
// tangential component; terms summed in the reference's order.  Triples per block as above.
__global__ __launch_bounds__(ST_NT) void force_v_kernel(const int* __restrict__ active, const DevBlock* __restrict__ blk,
                                                        const DevCase* __restrict__ cs, const double* __restrict__ ux,
                                                        const double* __restrict__ hux, const double* __restrict__ uy,
                                                        const double* __restrict__ huy, const double* __restrict__ uz,
                                                        const double* __restrict__ huz, const float* __restrict__ cut,
                                                        const uint8_t* __restrict__ cid8,
                                                        const uint32_t* __restrict__ flag, size_t capnc, int target,
                                                        double* __restrict__ part, size_t nslots) {
  __shared__ double red[ST_NT / 64];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const double mu = cs[B.case_id].mu;
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  const double dx = B.dx;
  const double* u[3] = {ux, uy, uz};
  const double* hu[3] = {hux, huy, huz};
  double fs[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + i + NB * (j + NB * k);
    const int ni[6] = {i - 1, i + 1, i, i, i, i}, nj[6] = {j, j, j - 1, j + 1, j, j}, nk[6] = {k, k, k, k, k - 1, k + 1};
    double up[3], un[3][6], mf[6];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      up[a] = u[a][g];
#pragma unroll
      for (int f = 0; f < 6; f++) un[a][f] = at(u[a], hu[a], slot, ni[f], nj[f], nk[f]);
    }
#pragma unroll
    for (int f = 0; f < 6; f++) {
      mf[f] = 0.0;
      if (cid8[f * capnc + g] == target) {
        const double w = 1.0 - 1.0 / (double)cut[f * capnc + g];
#pragma unroll
        for (int a = 0; a < 3; a++) un[a][f] = w * up[a];
        mf[f] = 1.0;
      }
    }
    if (!((flag[g] >> 6) & 1u))
#pragma unroll
      for (int f = 0; f < 6; f++) mf[f] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double acc = 0.0;
#pragma unroll
      for (int ax = 0; ax < 3; ax++) {
        const int fm = 2 * ax, fp = 2 * ax + 1;
        double lo, hi;
        if (ax == a) {
          lo = mu * (up[a] - un[a][fm]) / dx * mf[fm] * 2.0;
          hi = mu * (un[a][fp] - up[a]) / dx * mf[fp] * 2.0;
        } else {
          const double cross = (un[ax][2 * a + 1] - un[ax][2 * a]) * 0.5;
          lo = mu * ((up[a] - un[a][fm]) + cross) / dx * mf[fm];
          hi = mu * ((un[a][fp] - up[a]) + cross) / dx * mf[fp];
        }
        acc = ax == 0 ? lo - hi : acc + lo - hi;
      }
      fs[a] += acc * dx * dx;
    }
  }
  const double sx = block_sum<ST_NT>(fs[0], red), sy = block_sum<ST_NT>(fs[1], red), sz = block_sum<ST_NT>(fs[2], red);
  if (t == 0) {
    part[slot] = sx;
    part[nslots + slot] = sy;
    part[2 * nslots + slot] = sz;
  }
}

__global__ __launch_bounds__(ST_NT) void force_sum_kernel(const int* __restrict__ caseact,
                                                          const DevCase* __restrict__ cs,
                                                          const double* __restrict__ part, size_t nslots,
                                                          double* __restrict__ out) {
  __shared__ double red[ST_NT / 64];
  const DevCase& C = cs[caseact[blockIdx.x]];
  for (int q = 0; q < 3; q++) {
    const double v = case_sum<ST_NT>(part + q * nslots, C.slot_begin, C.nblk, red);
    if (threadIdx.x == 0) out[3 * blockIdx.x + q] = v;
  }
}

// ------------------------------------------------------------------ field statistics
// LocalScalar3D::CalcStats (LocalScalar3D.cpp:9-80, sf3d_calc_stats): sum, max, min, |max|,
// |min| over the interior cells of each case.  Block partials, then per-case combination in slot
// order (the sum with a fixed tree; max/min are order-free).
__device__ inline double block_op(double v, double* red, int op) {  // op 0 max, 1 min
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = op == 0 ? fmax(v, o) : fmin(v, o);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
  for (int q = 1; q < ST_NT / 64; q++) r = op == 0 ? fmax(r, red[q]) : fmin(r, red[q]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(ST_NT) void stats_kernel(const int* __restrict__ active, const double* __restrict__ x,
                                                      double* __restrict__ part, size_t nslots) {
  __shared__ double red[ST_NT / 64];
  const int slot = active[blockIdx.x];
  const size_t base = (size_t)slot * NC;
  double s = 0.0, mx = -DBL_MAX, mn = DBL_MAX, amx = 0.0, amn = DBL_MAX;
  for (int q = threadIdx.x; q < NC; q += ST_NT) {
    const double v = x[base + q];
    s += v;
    mx = fmax(mx, v);
    mn = fmin(mn, v);
    amx = fmax(amx, fabs(v));
    amn = fmin(amn, fabs(v));
  }
  const double ss = block_sum<ST_NT>(s, red);
  const double m0 = block_op(mx, red, 0), m1 = block_op(mn, red, 1), m2 = block_op(amx, red, 0),
               m3 = block_op(amn, red, 1);
  if (threadIdx.x == 0) {
    part[slot] = ss;
    part[nslots + slot] = m0;
    part[2 * nslots + slot] = m1;
    part[3 * nslots + slot] = m2;
    part[4 * nslots + slot] = m3;
  }
}

__global__ __launch_bounds__(ST_NT) void stats_sum_kernel(const int* __restrict__ caseact,
                                                          const DevCase* __restrict__ cs,
                                                          const double* __restrict__ part, size_t nslots,
                                                          double* __restrict__ out) {
  __shared__ double red[ST_NT / 64];
  const DevCase& C = cs[caseact[blockIdx.x]];
  const double sum = case_sum<ST_NT>(part, C.slot_begin, C.nblk, red);
  if (threadIdx.x != 0) return;
  double mx = -DBL_MAX, mn = DBL_MAX, amx = 0.0, amn = DBL_MAX;
  for (int b = 0; b < C.nblk; b++) {
    const size_t q = (size_t)C.slot_begin + b;
    mx = fmax(mx, part[nslots + q]);
    mn = fmin(mn, part[2 * nslots + q]);
    amx = fmax(amx, part[3 * nslots + q]);
    amn = fmin(amn, part[4 * nslots + q]);
  }
  double* o = out + 5 * blockIdx.x;
  o[0] = sum;
  o[1] = mx;
  o[2] = mn;
  o[3] = amx;
  o[4] = amn;
}

// ------------------------------------------------------------------ launchers
void launch_predict(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                    const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc) {
  if (nact) hipLaunchKernelGGL(predict_kernel, dim3(nact, 3), dim3(ST_NT), 0, sm, active, blk, cs, F, cut, flag, capnc);
}
void launch_abd_u(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                  const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc, int dir, double alpha) {
  if (nact)
    hipLaunchKernelGGL(abd_u_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, cs, F, cut, flag, capnc, dir, alpha);
}
void launch_p1(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
               const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc) {
  if (nact) hipLaunchKernelGGL(p1_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, cs, F, cut, flag, capnc);
}
void launch_ab_p(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                 const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc) {
  if (nact) hipLaunchKernelGGL(ab_p_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, cs, F, cut, flag, capnc);
}
void launch_refval(hipStream_t sm, const int* slots, int n, const DevBlock* blk, const StepFields& F, double xr,
                   double yr, double zr, double pr) {
  if (n) hipLaunchKernelGGL(refval_kernel, dim3(n), dim3(64), 0, sm, slots, blk, F, xr, yr, zr, pr);
}
void launch_matrix_bc(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                      const StepFields& F, int bcg) {
  if (nact) hipLaunchKernelGGL(matrix_bc_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, cs, F, bcg);
}
void launch_force_p(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase, const DevBlock* blk,
                    const DevCase* cs, const double* p, const double* hp, const uint8_t* cid8, const uint32_t* flag,
                    size_t capnc, int target, double* part, size_t nslots, double* out) {
  if (!nact || !ncase) return;
  hipLaunchKernelGGL(force_p_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, p, hp, cid8, flag, capnc, target,
                     part, nslots);
  hipLaunchKernelGGL(force_sum_kernel, dim3(ncase), dim3(ST_NT), 0, sm, caseact, cs, part, nslots, out);
}
void launch_force_v(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase, const DevBlock* blk,
                    const DevCase* cs, const double* const* u, const double* const* hu, const float* cut,
                    const uint8_t* cid8, const uint32_t* flag, size_t capnc, int target, double* part, size_t nslots,
                    double* out) {
  if (!nact || !ncase) return;
  hipLaunchKernelGGL(force_v_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, cs, u[0], hu[0], u[1], hu[1], u[2],
                     hu[2], cut, cid8, flag, capnc, target, part, nslots);
  hipLaunchKernelGGL(force_sum_kernel, dim3(ncase), dim3(ST_NT), 0, sm, caseact, cs, part, nslots, out);
}
void launch_stats(hipStream_t sm, const int* active, int nact, const int* caseact, int ncase, const DevCase* cs,
                  const double* x, double* part, size_t nslots, double* out) {
  if (!nact || !ncase) return;
  hipLaunchKernelGGL(stats_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, x, part, nslots);
  hipLaunchKernelGGL(stats_sum_kernel, dim3(ncase), dim3(ST_NT), 0, sm, caseact, cs, part, nslots, out);
}
void launch_corr(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const DevCase* cs,
                 const StepFields& F, const float* cut, const uint32_t* flag, size_t capnc) {
  if (nact) hipLaunchKernelGGL(corr_kernel, dim3(nact), dim3(ST_NT), 0, sm, active, blk, cs, F, cut, flag, capnc);
}

}  // namespace ffvhc
