// Dump2 (LocalScalar3D<real>::Dump2, LocalScalar3D.cpp:215-264): the reference fills the ghost
// cells of the field first (ImposeBoundaryCondition) and writes every block's padded array.  The
// device keeps the interior and one ghost layer, so the dump rebuilds all vc ghost layers on the
// host from the case's interiors, in the reference's order:
//   1. fine-to-coarse means into coarse blocks' ghosts (Scalar3DUpdater1.cpp:432-523 with
//      sup_copy_from_neighbor_f2c, sup.f90:122-179; Scalar3DUpdater5.h:650-763),
//   2. same-level copies (sup_copy_from_neighbor, sup.f90:18-52),
//   3. coarse-to-fine values into fine blocks' ghosts (Updater1: trilinear from the coarse interior,
//      sup_copy_from_neighbor_c2f + sup_get_intp_index, sup.f90:1-16, 54-120; Updater5:
//      f_ghost = f_inner + (c_in - c_ghost) / 2 layer by layer, Scalar3DUpdater5.h:765-885),
//   4. the outer-face BC over the whole padded face, faces in the order X_M X_P Y_M Y_P Z_M Z_P
//      (BlockScalar3D.h:174-254, bcx.f90: Dirichlet 5-147, Neumann 294-436, periodic 585-658; the
//      minus-side periodic routine is a no-op).
// Cells none of these writes (block edges and corners away from outer faces) keep the reference's
// earlier values there; here they are zero.
#include <cstdio>
#include <cstring>
#include <vector>

#include "ghost.h"
#include "runtime.h"

using namespace ffvhc;

namespace {

struct Padded {
  double* base;
  int g, np;
  size_t npad;
  // 0-based cell (i, j, k), each in [-g, NB + g), of block b
  double& at(int b, int i, int j, int k) const {
    return base[(size_t)b * npad + (size_t)(i + g) + (size_t)np * ((size_t)(j + g) + (size_t)np * (size_t)(k + g))];
  }
};

// cell of block-local face coordinates: normal layer l counted from face f into the block (l < 0:
// the ghost layers outside it, -1 adjacent) and face position (a, b)
void cell_of(int f, int l, int a, int b, int& i, int& j, int& k) {
  const int nrm = (f & 1) ? NB - 1 - l : l;
  switch (f >> 1) {
    case 0: i = nrm; j = a; k = b; break;
    case 1: i = a; j = nrm; k = b; break;
    default: i = a; j = b; k = nrm; break;
  }
}

void fill_f2c(const Padded& P, int cb, int f, int s, int fb, bool u5) {
  // coarse block cb, face f; fine neighbour fb holds quarter s; ghost layer l (0 adjacent) takes
  // the mean of the 2 x 2 x 2 fine cells of fine layers 2l and 2l + 1 from the shared face
  int oa, ob;
  sub_origins_ab(f, s, oa, ob);
  for (int l = 0; l < P.g; l++)
    for (int bb = 0; bb < NB / 2; bb++)
      for (int aa = 0; aa < NB / 2; aa++) {
        int i, j, k;
        cell_of(f, -1 - l, aa + (NB / 2) * oa, bb + (NB / 2) * ob, i, j, k);
        // lowest corner (in i, j, k) of the fine 2 x 2 x 2 group
        int fi, fj, fk, gi, gj, gk;
        cell_of(f ^ 1, 2 * l, 2 * aa, 2 * bb, fi, fj, fk);
        cell_of(f ^ 1, 2 * l + 1, 2 * aa + 1, 2 * bb + 1, gi, gj, gk);
        const int x = std::min(fi, gi), y = std::min(fj, gj), z = std::min(fk, gk);
        auto F = [&](int di, int dj, int dk) { return P.at(fb, x + di, y + dj, z + dk); };
        double v;
        if (!u5)
          v = F(0, 0, 0) + F(1, 0, 0) + F(0, 1, 0) + F(0, 0, 1) + F(0, 1, 1) + F(1, 0, 1) + F(1, 1, 0) + F(1, 1, 1);
        else
          v = F(0, 0, 0) + F(1, 0, 0) + F(0, 1, 0) + F(1, 1, 0) + F(0, 0, 1) + F(1, 0, 1) + F(0, 1, 1) + F(1, 1, 1);
        P.at(cb, i, j, k) = 0.125 * v;
      }
}

void fill_same(const Padded& P, int b, int f, int nb) {
  for (int l = 0; l < P.g; l++)
    for (int bb = 0; bb < NB; bb++)
      for (int aa = 0; aa < NB; aa++) {
        int i, j, k, si, sj, sk;
        cell_of(f, -1 - l, aa, bb, i, j, k);
        cell_of(f ^ 1, l, aa, bb, si, sj, sk);
        P.at(b, i, j, k) = P.at(nb, si, sj, sk);
      }
}

void fill_c2f(const Padded& P, int fb, int f, int s, int cb, bool u5) {
  int oa, ob;
  sub_origins_ab(f, s, oa, ob);
  const int nrm = f >> 1;
  int ax_a, ax_b;
  face_axes(f, ax_a, ax_b);
  for (int l = 0; l < P.g; l++)
    for (int bb = 0; bb < NB; bb++)
      for (int aa = 0; aa < NB; aa++) {
        int i, j, k;
        cell_of(f, -1 - l, aa, bb, i, j, k);
        if (u5) {
          // coarse cell across (face position a / 2 + 8 oa, b / 2 + 8 ob): its interior cell at the
          // face and the coarse ghost beside it (filled by the fine-to-coarse pass)
          const int ca = aa / 2 + (NB / 2) * oa, cbb = bb / 2 + (NB / 2) * ob;
          int ii, jj, kk, gi, gj, gk, pi, pj, pk;
          cell_of(f ^ 1, 0, ca, cbb, ii, jj, kk);
          cell_of(f ^ 1, -1, ca, cbb, gi, gj, gk);
          const double dp = (P.at(cb, ii, jj, kk) - P.at(cb, gi, gj, gk)) * 0.5;
          cell_of(f, -l, aa, bb, pi, pj, pk);  // the layer inside this one (l = 0: the interior)
          P.at(fb, i, j, k) = P.at(fb, pi, pj, pk) + dp;
        } else {
          // doubled 1-based index of the fine ghost cell in the coarse block's frame
          int d[3];
          const int gl = (f & 1) ? l : -1 - l;  // ghost offset beyond the face, in fine cells
          d[nrm] = (f & 1) ? 1 + gl : 2 * NB + 1 + gl;
          d[ax_a] = aa + 1 + NB * oa;
          d[ax_b] = bb + 1 + NB * ob;
          int ic[3];
          double w[3];
          for (int q = 0; q < 3; q++) intp_index(d[q], ic[q], w[q]);
          auto C = [&](int di, int dj, int dk) { return P.at(cb, ic[0] + di, ic[1] + dj, ic[2] + dk); };
          const double r = w[0], sw = w[1], t = w[2];
          const double a0 = (1.0 - r) * C(0, 0, 0) + r * C(1, 0, 0);
          const double a1 = (1.0 - r) * C(0, 1, 0) + r * C(1, 1, 0);
          const double a2 = (1.0 - r) * C(0, 0, 1) + r * C(1, 0, 1);
          const double a3 = (1.0 - r) * C(0, 1, 1) + r * C(1, 1, 1);
          const double c0 = (1.0 - sw) * a0 + sw * a1;
          const double c1 = (1.0 - sw) * a2 + sw * a3;
          P.at(fb, i, j, k) = (1.0 - t) * c0 + t * c1;
        }
      }
}

void fill_outer(const Padded& P, int b, int f, int type, double xc) {
  const int g = P.g;
  for (int bb = -g; bb < NB + g; bb++)
    for (int aa = -g; aa < NB + g; aa++) {
      auto A = [&](int l) -> double& {  // normal layer l from face f (0: the cell at the face)
        int i, j, k;
        cell_of(f, l, aa, bb, i, j, k);
        return P.at(b, i, j, k);
      };
      const bool plus = f & 1;
      if (type == BCT_DIRICHLET) {
        A(-1) = 2.0 * xc - A(0);
        if (g > 1) A(-2) = xc;
      } else if (type == BCT_NEUMANN) {
        // bc_x1_n: x(n+1) = x(n) + xc, x(n+2) = x(n) + 2 xc; bc_x3_n: x(0) = x(1) - xc, x(-1) = x(1) - 2 xc
        A(-1) = plus ? A(0) + xc : A(0) - xc;
        if (g > 1) A(-2) = plus ? A(0) + xc * 2.0 : A(0) - xc * 2.0;
      } else if (type == BCT_PERIODIC && plus) {
        // x(n+1) = x(1), x(n+2) = x(2), x(0) = x(n), x(-1) = x(n-1)
        A(-1) = A(NB - 1);
        if (g > 1) A(-2) = A(NB - 2);
        A(NB) = A(0);
        if (g > 1) A(NB + 1) = A(1);
      }
    }
}

}  // namespace

namespace ffvhc {

// Ghost layers 1..g of a case's padded host copy (interiors valid), as ImposeBoundaryCondition
// leaves them for BC group bcg and updater 1 or 5.
int host_impose_bc(ffvhc_ctx* c, int id, double* padded, int g, int bcg, int updater) {
  const CaseRec& r = c->cases[id];
  const int b0 = r.slot_begin, nb = r.nblk;
  Padded P{padded, g, NB + 2 * g, (size_t)(NB + 2 * g) * (NB + 2 * g) * (NB + 2 * g)};
  const bool u5 = updater == 5;
  auto local = [&](int slot) { return slot - b0; };
  for (int b = 0; b < nb; b++) {
    const DevBlock& B = c->hblk[b0 + b];
    for (int f = 0; f < 6; f++)
      if (B.ftype[f] == FACE_FINER)
        for (int s = 0; s < 4; s++) fill_f2c(P, b, f, s, local(B.nbr[f][s]), u5);
  }
  for (int b = 0; b < nb; b++) {
    const DevBlock& B = c->hblk[b0 + b];
    for (int f = 0; f < 6; f++)
      if (B.ftype[f] == FACE_SAME) fill_same(P, b, f, local(B.nbr[f][0]));
  }
  for (int b = 0; b < nb; b++) {
    const DevBlock& B = c->hblk[b0 + b];
    for (int f = 0; f < 6; f++)
      if (B.ftype[f] == FACE_COARSER) fill_c2f(P, b, f, B.sub[f], local(B.nbr[f][0]), u5);
  }
  const DevCase& C = c->hcase[id];
  for (int b = 0; b < nb; b++) {
    const DevBlock& B = c->hblk[b0 + b];
    for (int f = 0; f < 6; f++) {
      if (B.ftype[f] != FACE_OUTER) continue;
      const int type = C.bct[bcg][f];
      if (type < 0) continue;
      const double xc = type == BCT_NEUMANN ? C.bcv[bcg][f] * B.dx : C.bcv[bcg][f];
      fill_outer(P, b, f, type, xc);
    }
  }
  return 0;
}

}  // namespace ffvhc
