// Sphere as cut cells on the block octree (host, one-time setup).
//
// Replaces Cutlib's CalcCutInfo for an analytic sphere (Solver.cpp:530-682): along each axis
// direction of each cell, the cut position is the distance from the cell centre to the sphere
// surface in cell units, in [0, 1], stored as float32 (CutPos32Array); faces without a cut keep
// 1.0 and boundary id 0.  Then, as ModifyCut1/ModifyCut2 (Solver.cpp:684-1046): bstl_cutoff
// (bstl.f90:1306-1372), bstl_detect_zerocut (1557-1771), bstl_fill_holes_v1 (1773-2023) and the
// iterated bstl_fill_holes_v2 (2025-2275), each followed by a ghost update of the cut arrays;
// finally the phase flood fill from a seed point (FillCellFlag / InitPhase, Solver.cpp:1182-1287,
// 1404-1537): a cell joins the fluid when a neighbour is fluid and its own face toward that
// neighbour is not a phase boundary.
//
// Ghost updates (the reference's ImposeBoundaryCondition of the cut arrays): fine-to-coarse means
// and coarse-to-fine trilinear values at level interfaces (cut distances as reals, cut ids
// truncated to int as the Scalar3DUpdater1 template does), same-level neighbours copy the facing
// interior layers, outer faces take the Neumann-0 copy of the first interior layer
// (boundaryTypeNULL).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "../../include/ffvhc.h"
#include "ghost.h"
#include "octree.h"
#include "runtime.h"

namespace {

using ffvhc::Block;
using ffvhc::Octree;

// Runs f(b) for b in [0, nb) on up to 16 host threads.  Every pass below that runs this way
// writes only the padded arrays of its own block b (reading other blocks' interiors at most), so
// the blocks are independent and the result does not depend on the thread count.
template <class F>
void for_blocks(int nb, F f) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nt = (int)std::min<unsigned>(hw, (unsigned)std::max(nb, 1));
  if (nt <= 1) {
    for (int b = 0; b < nb; b++) f(b);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++)
    pool.emplace_back([&]() {
      for (int b = next.fetch_add(1); b < nb; b = next.fetch_add(1)) f(b);
    });
  for (auto& th : pool) th.join();
}

struct Arrays {
  int n, g, np, npad, nb;
  std::vector<double> c[6];  // [blk][npad]
  std::vector<int> id[6];
  size_t at(int b, int i, int j, int k) const {  // i, j, k in [-g, n+g)
    return (size_t)b * npad + (size_t)(i + g) + (size_t)np * ((size_t)(j + g) + (size_t)np * (size_t)(k + g));
  }
};

// padded coordinates of the cell at layer l (0 = adjacent) inside the block from face f, face
// position (a, b) along the two other axes in increasing axis order
void face_ijk(int f, int a, int b, int l, int n, int& i, int& j, int& k) {
  const int nrm = (f & 1) ? n - 1 - l : l;
  switch (f >> 1) {
    case 0: i = nrm; j = a; k = b; break;
    case 1: i = a; j = nrm; k = b; break;
    default: i = a; j = b; k = nrm; break;
  }
}

// Coarse/fine ghost updates of the cut arrays, as ImposeBoundaryCondition does them: the cut
// distances are LocalScalar3D<real> (Scalar3DUpdater1 with sup_copy_from_neighbor_f2c / _c2f,
// sup.f90:54-179), the cut ids LocalScalar3D<int> (the Scalar3DUpdater1 template,
// Scalar3DUpdater1.h:164-252 and 503-700: the same means and trilinear values computed in
// double and truncated to int on assignment).  f2c: a coarse block's ghost layers from its finer
// neighbours; c2f: a fine block's ghost layers from its coarser neighbour (interior values only).
void face_cell_l(int f, int l, int a, int b, int n, int& i, int& j, int& k) {
  // normal layer l counted from face f into the block (negative: ghost layers), position (a, b)
  face_ijk(f, a, b, 0, n, i, j, k);
  int* q = (f >> 1) == 0 ? &i : (f >> 1) == 1 ? &j : &k;
  *q = (f & 1) ? n - 1 - l : l;
}

void exchange_f2c(Arrays& A, const std::vector<Block>& blk) {
  const int n = A.n, g = A.g;
  for_blocks(A.nb, [&](int b) {
    const Block& B = blk[b];
    for (int f = 0; f < 6; f++) {
      if (B.ftype[f] != ffvhc::FACE_FINER) continue;
      for (int s = 0; s < 4; s++) {
        const int fb = B.nbr[f][s];
        int oa, ob;
        ffvhc::sub_origins_ab(f, s, oa, ob);
        for (int l = 0; l < g; l++)
          for (int bb = 0; bb < n / 2; bb++)
            for (int aa = 0; aa < n / 2; aa++) {
              int i, j, k, fi, fj, fk, gi, gj, gk;
              face_cell_l(f, -1 - l, aa + (n / 2) * oa, bb + (n / 2) * ob, n, i, j, k);
              face_cell_l(f ^ 1, 2 * l, 2 * aa, 2 * bb, n, fi, fj, fk);
              face_cell_l(f ^ 1, 2 * l + 1, 2 * aa + 1, 2 * bb + 1, n, gi, gj, gk);
              const int x = std::min(fi, gi), y = std::min(fj, gj), z = std::min(fk, gk);
              const size_t dst = A.at(b, i, j, k);
              auto at = [&](int di, int dj, int dk) { return A.at(fb, x + di, y + dj, z + dk); };
              const size_t q[8] = {at(0, 0, 0), at(1, 0, 0), at(0, 1, 0), at(0, 0, 1),
                                   at(0, 1, 1), at(1, 0, 1), at(1, 1, 0), at(1, 1, 1)};
              for (int d = 0; d < 6; d++) {
                double sc = 0.0;
                int si = 0;
                for (int m = 0; m < 8; m++) {  // sup order: nnn pnn npn nnp npp pnp ppn ppp
                  sc += A.c[d][q[m]];
                  si += A.id[d][q[m]];
                }
                A.c[d][dst] = 0.125 * sc;
                A.id[d][dst] = (int)(0.125 * (double)si);
              }
            }
      }
    }
  });
}

// linear-interpolation index of the Scalar3DUpdater1 template on a 0-based doubled index
void lin_index(int i, int n, int& I, double& r) {
  I = std::min(std::max(i / 2 - 1 + i % 2, 0), n - 2);
  r = -0.25 + 0.5 * i - (double)I;
}

void exchange_c2f(Arrays& A, const std::vector<Block>& blk) {
  const int n = A.n, g = A.g;
  for_blocks(A.nb, [&](int b) {
    const Block& B = blk[b];
    for (int f = 0; f < 6; f++) {
      if (B.ftype[f] != ffvhc::FACE_COARSER) continue;
      const int cb = B.nbr[f][0];
      int oa, ob, ax_a, ax_b;
      ffvhc::sub_origins_ab(f, B.sub[f], oa, ob);
      ffvhc::face_axes(f, ax_a, ax_b);
      const int nrm = f >> 1;
      for (int l = 0; l < g; l++)
        for (int bb = 0; bb < n; bb++)
          for (int aa = 0; aa < n; aa++) {
            int i, j, k;
            face_cell_l(f, -1 - l, aa, bb, n, i, j, k);
            // 0-based doubled index of the ghost cell in the coarse block's frame
            int dd[3];
            dd[nrm] = (f & 1) ? l : 2 * n - 1 - l;
            dd[ax_a] = aa + n * oa;
            dd[ax_b] = bb + n * ob;
            int I[3];
            double w[3];
            for (int q = 0; q < 3; q++) lin_index(dd[q], n, I[q], w[q]);
            const double r = w[0], sw = w[1], t = w[2];
            const size_t dst = A.at(b, i, j, k);
            auto at = [&](int di, int dj, int dk) { return A.at(cb, I[0] + di, I[1] + dj, I[2] + dk); };
            const size_t c000 = at(0, 0, 0), c100 = at(1, 0, 0), c010 = at(0, 1, 0), c110 = at(1, 1, 0);
            const size_t c001 = at(0, 0, 1), c101 = at(1, 0, 1), c011 = at(0, 1, 1), c111 = at(1, 1, 1);
            for (int d = 0; d < 6; d++) {
              auto tri = [&](auto V) {
                return (1.0 - t) * ((1.0 - sw) * ((1.0 - r) * V(c000) + r * V(c100)) +
                                    sw * ((1.0 - r) * V(c010) + r * V(c110))) +
                       t * ((1.0 - sw) * ((1.0 - r) * V(c001) + r * V(c101)) +
                            sw * ((1.0 - r) * V(c011) + r * V(c111)));
              };
              const std::vector<double>& cv = A.c[d];
              const std::vector<int>& iv = A.id[d];
              A.c[d][dst] = tri([&](size_t m) { return cv[m]; });
              A.id[d][dst] = (int)tri([&](size_t m) { return (double)iv[m]; });
            }
          }
    }
  });
}

void exchange(Arrays& A, const std::vector<Block>& blk) {
  const int n = A.n, g = A.g;
  exchange_f2c(A, blk);
  for_blocks(A.nb, [&](int b) {
    for (int f = 0; f < 6; f++) {
      const Block& B = blk[b];
      if (B.ftype[f] != ffvhc::FACE_SAME && B.ftype[f] != ffvhc::FACE_OUTER) continue;
      for (int l = 1; l <= g; l++)
        for (int bb = 0; bb < n; bb++)
          for (int aa = 0; aa < n; aa++) {
            int gi, gj, gk, si, sj, sk;
            face_ijk(f, aa, bb, -l, n, gi, gj, gk);  // ghost layer l outside face f
            size_t dst = A.at(b, gi, gj, gk), src;
            if (B.ftype[f] == ffvhc::FACE_SAME) {
              face_ijk(f ^ 1, aa, bb, l - 1, n, si, sj, sk);
              src = A.at(B.nbr[f][0], si, sj, sk);
            } else {
              face_ijk(f, aa, bb, 0, n, si, sj, sk);
              src = A.at(b, si, sj, sk);
            }
            for (int d = 0; d < 6; d++) {
              A.c[d][dst] = A.c[d][src];
              A.id[d][dst] = A.id[d][src];
            }
          }
    }
  });
  exchange_c2f(A, blk);
}

}  // namespace

extern "C" int ffvhc_geometry_sphere(const ffvhc_tree* tree, const ffvhc_sphere_spec* S, int g, double* c, int* cid,
                                     int* pid, long long* counts) {
  using ffvhc::fail;
  if (!tree || !S || !pid) return fail(-1, "null argument");
  if (g < 1) return fail(-1, "g must be >= 1");
  if (!(S->radius > 0.0)) return fail(-1, "radius must be positive");
  if (S->cut_id < 1 || S->cut_id > 31) return fail(-1, "cut_id must be in 1..31");
  const Octree& T = *tree->tree;
  const std::vector<Block>& blk = T.blocks();
  Arrays A;
  A.n = T.spec().n;
  A.g = g;
  A.np = A.n + 2 * g;
  A.npad = A.np * A.np * A.np;
  A.nb = (int)blk.size();
  for (int d = 0; d < 6; d++) {
    A.c[d].assign((size_t)A.nb * A.npad, 1.0);
    A.id[d].assign((size_t)A.nb * A.npad, 0);
  }
  const int n = A.n;
  const double R2 = S->radius * S->radius;
  // analytic cuts for every padded cell (a block whose padded box stays farther than the radius
  // from the centre has none)
  for_blocks(A.nb, [&](int b) {
    const Block& B = blk[b];
    double d2 = 0.0;
    for (int a = 0; a < 3; a++) {
      const double lo = B.org[a] - g * B.dx, hi = B.org[a] + (n + g) * B.dx;
      const double e = S->center[a] < lo ? lo - S->center[a] : S->center[a] > hi ? S->center[a] - hi : 0.0;
      d2 += e * e;
    }
    if (d2 > R2) return;
    for (int k = -g; k < n + g; k++)
      for (int j = -g; j < n + g; j++)
        for (int i = -g; i < n + g; i++) {
          const double xc[3] = {B.org[0] + (i + 0.5) * B.dx, B.org[1] + (j + 0.5) * B.dx, B.org[2] + (k + 0.5) * B.dx};
          const size_t m = A.at(b, i, j, k);
          for (int f = 0; f < 6; f++) {
            const int ax = f >> 1;
            const double s = (f & 1) ? 1.0 : -1.0;
            double q = 0.0;
            for (int a = 0; a < 3; a++)
              if (a != ax) q += (xc[a] - S->center[a]) * (xc[a] - S->center[a]);
            if (R2 - q < 0.0) continue;
            const double r = std::sqrt(R2 - q), delta = xc[ax] - S->center[ax];
            // s t dx = -delta +- r  ->  t = s (-delta +- r) / dx; smallest t in [0, 1]
            double best = 2.0;
            for (double root : {-delta - r, -delta + r}) {
              const double t = s * root / B.dx;
              if (t >= 0.0 && t <= 1.0 && t < best) best = t;
            }
            if (best <= 1.0) {
              A.c[f][m] = (double)(float)best;
              A.id[f][m] = S->cut_id;
            }
          }
        }
  });
  exchange(A, blk);
  // bstl_cutoff (ShapeApproximation/Method "cut", Cutoff > 0)
  if (S->cutoff > 0.0) {
    for_blocks(A.nb, [&](int b) {
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            const size_t m = A.at(b, i, j, k);
            for (int d = 0; d < 6; d++) A.c[d][m] = std::max(S->cutoff, A.c[d][m]);
          }
    });
    exchange(A, blk);
  }
  // bstl_detect_zerocut
  std::atomic<long long> nzero{0};
  for_blocks(A.nb, [&](int b) {
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          const size_t m = A.at(b, i, j, k);
          double cmin = 1.0;
          int idmin = 0;
          for (int d = 0; d < 6; d++)
            if (cmin > A.c[d][m]) {
              cmin = A.c[d][m];
              idmin = A.id[d][m];
            }
          if (cmin == 0.0) {
            for (int d = 0; d < 6; d++) A.c[d][m] = 1.0;
            const int nbr[6][4] = {{i - 1, j, k, 1}, {i + 1, j, k, 0}, {i, j - 1, k, 3},
                                   {i, j + 1, k, 2}, {i, j, k - 1, 5}, {i, j, k + 1, 4}};
            for (auto& q : nbr) {
              const size_t mm = A.at(b, q[0], q[1], q[2]);
              A.c[q[3]][mm] = 1.0;
              A.id[q[3]][mm] = idmin;
            }
            nzero++;
          }
        }
  });
  exchange(A, blk);
  // bstl_fill_holes_v1 / v2: close an open face whose four side faces are cut (v1), or either
  // pair of them (v2); one sweep in (k, j, i) order per pass
  auto fill_pass = [&](bool v2, bool close) {
    std::atomic<long long> total{0};
    for_blocks(A.nb, [&](int b) {
      long long cnt = 0;
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            const size_t p = A.at(b, i, j, k), w = A.at(b, i - 1, j, k), e = A.at(b, i + 1, j, k);
            const size_t s = A.at(b, i, j - 1, k), nn = A.at(b, i, j + 1, k), bt = A.at(b, i, j, k - 1),
                         t = A.at(b, i, j, k + 1);
            auto C = [&](int d, size_t q) { return A.id[d][q] != 0; };
            auto hole = [&](bool a1, bool a2, bool a3, bool a4) { return v2 ? ((a1 && a2) || (a3 && a4)) : (a1 && a2 && a3 && a4); };
            auto closef = [&](int d, size_t q, int d2, size_t q2) {
              if (close) {
                A.id[d][q] = 1;
                A.id[d2][q2] = 1;
                A.c[d][q] = 0.5;
                A.c[d2][q2] = 0.5;
              }
              cnt++;
            };
            if (A.id[0][p] == 0 && hole(C(0, s) || C(2, p) || C(2, w), C(0, nn) || C(3, p) || C(3, w),
                                        C(0, bt) || C(4, p) || C(4, w), C(0, t) || C(5, p) || C(5, w)))
              closef(0, p, 1, w);
            if (A.id[1][p] == 0 && hole(C(1, s) || C(2, p) || C(2, e), C(1, nn) || C(3, p) || C(3, e),
                                        C(1, bt) || C(4, p) || C(4, e), C(1, t) || C(5, p) || C(5, e)))
              closef(1, p, 0, e);
            if (A.id[2][p] == 0 && hole(C(2, w) || C(0, p) || C(0, s), C(2, e) || C(1, p) || C(1, s),
                                        C(2, bt) || C(4, p) || C(4, s), C(2, t) || C(5, p) || C(5, s)))
              closef(2, p, 3, s);
            if (A.id[3][p] == 0 && hole(C(3, w) || C(0, p) || C(0, nn), C(3, e) || C(1, p) || C(1, nn),
                                        C(3, bt) || C(4, p) || C(4, nn), C(3, t) || C(5, p) || C(5, nn)))
              closef(3, p, 2, nn);
            if (A.id[4][p] == 0 && hole(C(4, w) || C(0, p) || C(0, bt), C(4, e) || C(1, p) || C(1, bt),
                                        C(4, s) || C(2, p) || C(2, bt), C(4, nn) || C(3, p) || C(3, bt)))
              closef(4, p, 5, bt);
            if (A.id[5][p] == 0 && hole(C(5, w) || C(0, p) || C(0, t), C(5, e) || C(1, p) || C(1, t),
                                        C(5, s) || C(2, p) || C(2, t), C(5, nn) || C(3, p) || C(3, t)))
              closef(5, p, 4, t);
          }
      total += cnt;
    });
    exchange(A, blk);
    return total.load();
  };
  const long long nholes1 = fill_pass(false, S->hole_filling[0] != 0);
  long long nholes2 = 0;
  for (int it = 0; it < 1000; it++) {
    const long long k2 = fill_pass(true, S->hole_filling[1] != 0);
    nholes2 += k2;
    if (k2 == 0 || !S->hole_filling[1]) break;
  }
  // phase flood fill from the seed (faces with the sphere's id are phase boundaries)
  std::vector<uint8_t> fluid((size_t)A.nb * n * n * n, 0);
  auto cell = [&](int b, int i, int j, int k) { return ((size_t)b * n + k) * n * n + (size_t)j * n + i; };
  std::deque<std::array<int, 4>> q;
  for (int b = 0; b < A.nb; b++) {
    const Block& B = blk[b];
    int ijk[3];
    bool in = true;
    for (int a = 0; a < 3; a++) {
      const double rel = (S->seed[a] - B.org[a]) / B.dx;
      ijk[a] = (int)std::floor(rel);
      if (ijk[a] < 0 || ijk[a] >= n) in = false;
    }
    if (in) {
      fluid[cell(b, ijk[0], ijk[1], ijk[2])] = 1;
      q.push_back({b, ijk[0], ijk[1], ijk[2]});
      break;
    }
  }
  if (q.empty()) return fail(-1, "the seed point lies outside the grid");
  auto blocked = [&](int b, int i, int j, int k, int f) { return A.id[f][A.at(b, i, j, k)] == S->cut_id; };
  while (!q.empty()) {
    const auto cur = q.front();
    q.pop_front();
    const int b = cur[0];
    const Block& B = blk[b];
    for (int f = 0; f < 6; f++) {
      int p[3] = {cur[1], cur[2], cur[3]};
      const int ax = f >> 1;
      p[ax] += (f & 1) ? 1 : -1;
      // the neighbour enters through its own face toward the current cell (f ^ 1)
      auto visit = [&](int nb_, int i, int j, int k) {
        const size_t id = cell(nb_, i, j, k);
        if (fluid[id] || blocked(nb_, i, j, k, f ^ 1)) return;
        fluid[id] = 1;
        q.push_back({nb_, i, j, k});
      };
      if (p[ax] >= 0 && p[ax] < n) {
        visit(b, p[0], p[1], p[2]);
        continue;
      }
      const int ft = B.ftype[f];
      if (ft == ffvhc::FACE_OUTER) continue;
      // position of the neighbouring cell centre, then the cell of the neighbour block there
      double x[3];
      for (int a = 0; a < 3; a++) x[a] = B.org[a] + (cur[1 + a] + 0.5) * B.dx;
      x[ax] += ((f & 1) ? 1.0 : -1.0) * B.dx;
      auto locate = [&](int nb_) {
        const Block& N = blk[nb_];
        int q3[3];
        for (int a = 0; a < 3; a++) {
          q3[a] = (int)std::floor((x[a] - N.org[a]) / N.dx);
          q3[a] = std::min(std::max(q3[a], 0), n - 1);
        }
        visit(nb_, q3[0], q3[1], q3[2]);
      };
      if (ft == ffvhc::FACE_SAME || ft == ffvhc::FACE_COARSER) {
        locate(B.nbr[f][0]);
      } else {  // finer neighbours: the two fine cells of each quarter facing this cell
        for (int s = 0; s < 4; s++) {
          const Block& N = blk[B.nbr[f][s]];
          for (int d1 = 0; d1 < 2; d1++)
            for (int d2 = 0; d2 < 2; d2++) {
              double y[3] = {x[0], x[1], x[2]};
              int o = 0;
              for (int a = 0; a < 3; a++) {
                if (a == ax) {
                  y[a] = B.org[a] + (cur[1 + a] + 0.5) * B.dx + ((f & 1) ? 0.75 : -0.75) * B.dx;
                } else {
                  y[a] += ((o++ == 0 ? d1 : d2) ? 0.25 : -0.25) * B.dx;
                }
              }
              int q3[3];
              bool inside = true;
              for (int a = 0; a < 3; a++) {
                q3[a] = (int)std::floor((y[a] - N.org[a]) / N.dx);
                if (q3[a] < 0 || q3[a] >= n) inside = false;
              }
              if (inside) visit(B.nbr[f][s], q3[0], q3[1], q3[2]);
            }
        }
      }
    }
  }
  std::atomic<long long> nf{0}, ns{0};
  const size_t np = (size_t)A.npad;
  for_blocks(A.nb, [&](int b) {
    long long f1 = 0, s1 = 0;
    for (int k = -g; k < n + g; k++)
      for (int j = -g; j < n + g; j++)
        for (int i = -g; i < n + g; i++) {
          const size_t m = A.at(b, i, j, k);
          const bool inside = i >= 0 && i < n && j >= 0 && j < n && k >= 0 && k < n;
          int v = 0;
          if (inside) {
            v = fluid[cell(b, i, j, k)] ? 1 : 0;
            if (v) f1++;
            else s1++;
          }
          pid[m] = v;
          for (int d = 0; d < 6; d++) {
            if (c) c[(size_t)d * A.nb * np + m] = A.c[d][m];
            if (cid) cid[(size_t)d * A.nb * np + m] = A.id[d][m];
          }
        }
    nf += f1;
    ns += s1;
  });
  if (counts) {
    counts[0] = nf;
    counts[1] = ns;
    counts[2] = nholes1;
    counts[3] = nholes2;
  }
  (void)nzero;
  return 0;
}
