// Block octree of equal-size Cartesian blocks: the part of BCMTools that FFV-HC's time
// step depends on (BCMOctree, NeighborInfo, block ordering), rebuilt from its observable
// contract (Solver.cpp:219-387, Scalar3DUpdater1.cpp:264-343/449-505).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "common.h"

namespace ffvhc {

// A refinement region: blocks whose (margin-expanded) box intersects the region are
// refined until they reach `level`.
struct Region {
  enum Kind { BALL = 0, SHELL = 1, BOX = 2 } kind;
  double a[3];   // BALL/SHELL centre, BOX min corner
  double b[3];   // BOX max corner
  double r;      // BALL/SHELL radius
  double dr;     // SHELL thickness parameter (SphereDivider2: shell [r-dr/2, r+dr])
  int level;
};

struct OctreeSpec {
  int root[3] = {1, 1, 1};
  double origin[3] = {0, 0, 0};
  double root_length = 1.0;
  int periodic[3] = {0, 0, 0};
  int n = 16;            // cells per block edge
  int min_level = 0;     // every block is refined to at least this level
  double margin = 0.0;   // search-region margin in cells of the finest region level
  int ordering = 1;      // 0 Z (Morton), 1 Hilbert (FFVConfig.cpp default "Hilbert")
  std::vector<Region> regions;
};

struct Block {
  int level;
  int64_t ix, iy, iz;  // block coordinates at its level over the whole root grid
  double org[3];
  double size;          // block edge length
  double dx;            // cell size
  int ftype[6];
  int nbr[6][4];
  int sub[6];
};

class Octree {
 public:
  explicit Octree(const OctreeSpec& spec);
  const std::vector<Block>& blocks() const { return blocks_; }
  const OctreeSpec& spec() const { return spec_; }
  int max_level() const { return max_level_; }
  int64_t num_cells() const { return (int64_t)blocks_.size() * spec_.n * spec_.n * spec_.n; }

 private:
  bool refine(int level, int64_t x, int64_t y, int64_t z) const;
  void balance();
  void order();
  void link();
  OctreeSpec spec_;
  int max_level_ = 0;
  std::vector<Block> blocks_;
};

// Hilbert index of cell (x,y,z) on a 2^bits lattice (Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004: AxestoTranspose, then bit interleave).
uint64_t hilbert_index(uint32_t x, uint32_t y, uint32_t z, int bits);
uint64_t morton_index(uint32_t x, uint32_t y, uint32_t z, int bits);

}  // namespace ffvhc
