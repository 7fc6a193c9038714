// Phase flood fill on the device (FillCellFlag / InitPhase, Solver.cpp:1182-1287 and 1404-1537):
// a cell is fluid when it can be reached from the seed cell through faces that are not phase
// boundaries; a move into a cell is blocked when the entered cell's own face toward the cell it
// comes from carries the boundary id.
//
// The reference repeats block sweeps in which a cell takes the seed's id when a neighbour (a ghost
// across a block face) holds it and the cell's own face toward that neighbour is not a boundary,
// refreshing the CellFlag ghosts by Scalar3DUpdater6 between sweeps: a coarse ghost takes the id
// when any of the 2 x 2 fine cells along the interface holds it (interpolateF2C_X_M ..
// Scalar3DUpdater6.h:286-376), a fine ghost when the coarse cell holding it does
// (interpolateC2F, 378-387).
//
// Labels: one 16-bit mask per (slot, column c = i + 16 j), bit k = cell (i, j, k) reached.  A pass
// runs one 256-thread workgroup per block (thread = column): it pulls the labels of the
// neighbouring blocks' facing cells (as the previous passes left them) by those same rules --
// same level: the facing cell; from a coarser neighbour: the coarse cell whose face footprint
// holds the fine cell; from finer neighbours: the 2 x 2 fine cells under the coarse cell -- then
// closes the block locally in LDS (x and y through the neighbouring columns, k through the
// column's own bits) and stores the labels.  Labels only ever go from 0 to 1, so repeating passes
// until one changes nothing gives the same fixed point as the reference's sweeps (and as the host
// geometry's breadth-first search, which moves by the same rules).
#include "kcommon.h"
#include "kernels.h"

namespace ffvhc {

constexpr int PH_NT = NF;  // one thread per (i, j) column

__device__ inline bool lab_bit(const uint16_t* lab, int slot, int i, int j, int k) {
  return (lab[(size_t)slot * NF + i + NB * j] >> k) & 1u;
}

__global__ __launch_bounds__(PH_NT) void phase_pass_kernel(const int* __restrict__ active, int nact,
                                                           const DevBlock* __restrict__ blk,
                                                           const uint8_t* __restrict__ cid8, size_t capnc, int target,
                                                           uint16_t* __restrict__ lab, int* __restrict__ changed) {
  __shared__ uint16_t L[NF];
  const int slot = active[blockIdx.x];
  const DevBlock& B = blk[slot];
  const int t = threadIdx.x, i = t & 15, j = t >> 4;
  // ent[f] bit k: cell (i, j, k) may be entered through its face f
  uint16_t ent[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    uint16_t m = 0;
    const uint8_t* idf = cid8 + (size_t)f * capnc + (size_t)slot * NC + i + NB * j;
#pragma unroll
    for (int k = 0; k < NB; k++)
      if (idf[NF * k] != target) m |= (uint16_t)(1u << k);
    ent[f] = m;
  }
  const uint16_t old = lab[(size_t)slot * NF + t];
  uint16_t mine = old;
  // pulls across the block's faces, cell by cell of this column's face cells
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int ft = B.ftype[f];
    if (ft == FACE_OUTER) continue;
    const int ax = f >> 1;
    const int here = (ax == 0) ? i : (ax == 1) ? j : -1;
    if (ax < 2 && here != ((f & 1) ? NB - 1 : 0)) continue;  // this column does not touch face f
    for (int kk = 0; kk < NB; kk++) {
      const int k = (ax == 2) ? ((f & 1) ? NB - 1 : 0) : kk;
      if (ax == 2 && kk > 0) break;  // a z face holds one cell of the column
      if (!((ent[f] >> k) & 1u) || ((mine >> k) & 1u)) continue;
      // face position (a, b) of cell (i, j, k) on face f
      const int a = ax == 0 ? j : i, b = ax == 2 ? j : k;
      bool in = false;
      if (ft == FACE_SAME) {
        int si, sj, sk;
        face_cell(f ^ 1, a, b, 0, si, sj, sk);
        in = lab_bit(lab, B.nbr[f][0], si, sj, sk);
      } else if (ft == FACE_COARSER) {
        int ca, cb, si, sj, sk;
        c2f_coarse_ab(f, B.sub[f], a, b, ca, cb);
        face_cell(f ^ 1, ca, cb, 0, si, sj, sk);
        in = lab_bit(lab, B.nbr[f][0], si, sj, sk);
      } else {  // FACE_FINER: the 2 x 2 fine cells of quarter s under this coarse cell
        const int s = ghost_source(B, slot, f, a, b);
        for (int d = 0; d < 4 && !in; d++) {
          const int fa = 2 * (a % (NB / 2)) + (d & 1), fb = 2 * (b % (NB / 2)) + (d >> 1);
          int si, sj, sk;
          face_cell(f ^ 1, fa, fb, 0, si, sj, sk);
          in = lab_bit(lab, s, si, sj, sk);
        }
      }
      if (in) mine |= (uint16_t)(1u << k);
    }
  }
  // local closure in LDS
  L[t] = mine;
  __syncthreads();
  for (;;) {
    uint16_t v = mine;
    if (i > 0) v |= L[t - 1] & ent[0];
    if (i < NB - 1) v |= L[t + 1] & ent[1];
    if (j > 0) v |= L[t - NB] & ent[2];
    if (j < NB - 1) v |= L[t + NB] & ent[3];
    for (;;) {  // along the column: from k - 1 through face 4, from k + 1 through face 5
      const uint16_t w = v | (uint16_t)((v << 1) & ent[4]) | (uint16_t)((v >> 1) & ent[5]);
      if (w == v) break;
      v = w;
    }
    const int grew = v != mine;
    __syncthreads();
    L[t] = v;
    mine = v;
    if (!__syncthreads_or(grew)) break;
  }
  lab[(size_t)slot * NF + t] = mine;
  if (mine != old) *changed = 1;  // the same value from every writer
}

// fluid bit (6) of the cell flags from the labels; per block the number of fluid cells and
// whether the block holds any cut or solid cell (DevBlock::flags)
__global__ __launch_bounds__(PH_NT) void phase_store_kernel(const int* __restrict__ active,
                                                            const uint16_t* __restrict__ lab,
                                                            uint32_t* __restrict__ flag, int* __restrict__ nfluid,
                                                            int* __restrict__ irregular) {
  __shared__ int red[PH_NT / 64][2];
  const int slot = active[blockIdx.x];
  const int t = threadIdx.x;
  const uint16_t m = lab[(size_t)slot * NF + t];
  int cnt = 0, irr = 0;
  for (int k = 0; k < NB; k++) {
    const size_t g = (size_t)slot * NC + t + NF * k;
    const uint32_t fl = (flag[g] & ~(1u << 6)) | (((m >> k) & 1u) << 6);
    flag[g] = fl;
    cnt += (m >> k) & 1u;
    irr |= fl != (1u << 6);
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off, 64);
    irr |= __shfl_xor(irr, off, 64);
  }
  if ((t & 63) == 0) {
    red[t >> 6][0] = cnt;
    red[t >> 6][1] = irr;
  }
  __syncthreads();
  if (t == 0) {
    int c = 0, r = 0;
    for (int w = 0; w < PH_NT / 64; w++) {
      c += red[w][0];
      r |= red[w][1];
    }
    nfluid[slot] = c;
    irregular[slot] = r;
  }
}

void launch_phase_pass(hipStream_t sm, const int* active, int nact, const DevBlock* blk, const uint8_t* cid8,
                       size_t capnc, int target, uint16_t* lab, int* changed) {
  if (nact)
    hipLaunchKernelGGL(phase_pass_kernel, dim3(nact), dim3(PH_NT), 0, sm, active, nact, blk, cid8, capnc, target, lab,
                       changed);
}

void launch_phase_store(hipStream_t sm, const int* active, int nact, const uint16_t* lab, uint32_t* flag,
                        int* nfluid, int* irregular) {
  if (nact) hipLaunchKernelGGL(phase_store_kernel, dim3(nact), dim3(PH_NT), 0, sm, active, lab, flag, nfluid, irregular);
}

}  // namespace ffvhc
