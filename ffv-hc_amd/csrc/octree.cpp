// Block octree: refinement, 2:1 face balance, Z / Hilbert ordering and neighbour tables.
//
// The refinement predicates follow SphereDivider2.h:75-155 (shell: refine unless the box is
// entirely inside radius r-dr/2 or entirely outside r+dr) and the bounding-box lists of
// SphereDivider3.h:72-112 / BlockBoundingBox.h:80-87.  The NeighborInfo contract follows the
// index setups of Scalar3DUpdater1.cpp:264-343 and 449-505: a face sees the outer boundary,
// one block of the same level, one coarser block (this block is quarter `sub` of its face) or
// four finer blocks indexed by subface; subface s has tangential origins (s&1, s>>1) on the
// axes (y,z) of X faces, (z,x) of Y faces and (x,y) of Z faces.
#include "octree.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace ffvhc {

namespace {

struct Key {
  int level;
  int64_t x, y, z;
  bool operator==(const Key& o) const { return level == o.level && x == o.x && y == o.y && z == o.z; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = (uint64_t)k.level * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)k.x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.y + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.z + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// tangential axes of a face: (axis of origin0, axis of origin1)
inline void tangential(int face, int& t0, int& t1) {
  switch (face / 2) {
    case 0: t0 = 1; t1 = 2; break;  // X faces: (y, z)
    case 1: t0 = 2; t1 = 0; break;  // Y faces: (z, x)
    default: t0 = 0; t1 = 1; break; // Z faces: (x, y)
  }
}

}  // namespace

uint64_t hilbert_index(uint32_t x, uint32_t y, uint32_t z, int bits) {
  if (bits <= 0) return 0;
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1);
  // Skilling 2004, AxestoTranspose: inverse undo excess work, then Gray encode.
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    uint32_t P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t h = 0;
  for (int bit = bits - 1; bit >= 0; bit--)
    for (int i = 0; i < 3; i++) h = (h << 1) | ((X[i] >> bit) & 1u);
  return h;
}

uint64_t morton_index(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint64_t h = 0;
  for (int bit = bits - 1; bit >= 0; bit--)
    h = (h << 3) | (uint64_t)((((z >> bit) & 1u) << 2) | (((y >> bit) & 1u) << 1) | ((x >> bit) & 1u));
  return h;
}

Octree::Octree(const OctreeSpec& spec) : spec_(spec) {
  for (int d = 0; d < 3; d++)
    if (spec_.root[d] < 1) throw std::invalid_argument("root grid must be >= 1 in every direction");
  if (spec_.n < 2 || (spec_.n & 1)) throw std::invalid_argument("cells per block must be even and >= 2");
  int top = spec_.min_level;
  for (const Region& r : spec_.regions) top = std::max(top, r.level);
  if (top > 20) throw std::invalid_argument("refinement level too deep");
  // depth-first refinement from every root
  std::vector<Key> stack, leaves;
  for (int z = 0; z < spec_.root[2]; z++)
    for (int y = 0; y < spec_.root[1]; y++)
      for (int x = 0; x < spec_.root[0]; x++) stack.push_back({0, x, y, z});
  while (!stack.empty()) {
    Key k = stack.back();
    stack.pop_back();
    if (refine(k.level, k.x, k.y, k.z)) {
      for (int c = 0; c < 8; c++)
        stack.push_back({k.level + 1, 2 * k.x + (c & 1), 2 * k.y + ((c >> 1) & 1), 2 * k.z + ((c >> 2) & 1)});
    } else {
      leaves.push_back(k);
    }
  }
  blocks_.clear();
  for (const Key& k : leaves) {
    Block b{};
    b.level = k.level;
    b.ix = k.x;
    b.iy = k.y;
    b.iz = k.z;
    blocks_.push_back(b);
  }
  balance();
  order();
  link();
}

bool Octree::refine(int level, int64_t x, int64_t y, int64_t z) const {
  if (level < spec_.min_level) return true;
  const double size = spec_.root_length / double(int64_t(1) << level);
  const double lo[3] = {spec_.origin[0] + x * size, spec_.origin[1] + y * size, spec_.origin[2] + z * size};
  for (const Region& r : spec_.regions) {
    if (r.level <= level) continue;
    const double m = spec_.margin * spec_.root_length / double(int64_t(1) << r.level) / spec_.n;
    double mn[3], mx[3];
    for (int d = 0; d < 3; d++) {
      mn[d] = lo[d] - m;
      mx[d] = lo[d] + size + m;
    }
    if (r.kind == Region::BOX) {
      bool hit = true;
      for (int d = 0; d < 3; d++) hit = hit && (mn[d] < r.b[d]) && (mx[d] > r.a[d]);
      if (hit) return true;
      continue;
    }
    // nearest / farthest squared distance from the centre to the box
    double near2 = 0.0, far2 = 0.0;
    for (int d = 0; d < 3; d++) {
      double c = r.a[d];
      double dn = (c < mn[d]) ? mn[d] - c : (c > mx[d] ? c - mx[d] : 0.0);
      double df = std::max(std::fabs(mn[d] - c), std::fabs(mx[d] - c));
      near2 += dn * dn;
      far2 += df * df;
    }
    if (r.kind == Region::BALL) {
      if (near2 < r.r * r.r) return true;
    } else {  // SHELL, SphereDivider2::checkInner / checkOuter
      double rin = r.r - 0.5 * r.dr, rout = r.r + 1.0 * r.dr;
      bool inner = far2 < rin * rin;
      bool outer = near2 >= rout * rout;
      if (!inner && !outer) return true;
    }
  }
  return false;
}

void Octree::balance() {
  // Refine any leaf that has a face neighbour two or more levels finer, until 2:1 holds.
  std::unordered_set<Key, KeyHash> set;
  for (const Block& b : blocks_) set.insert({b.level, b.ix, b.iy, b.iz});
  bool changed = true;
  while (changed) {
    changed = false;
    std::vector<Key> split;
    for (const Key& k : set) {
      for (int f = 0; f < 6; f++) {
        int axis = f / 2;
        int64_t p[3] = {k.x, k.y, k.z};
        p[axis] += (f & 1) ? 1 : -1;
        int64_t ext = (int64_t)spec_.root[axis] << k.level;
        if (p[axis] < 0 || p[axis] >= ext) {
          if (!spec_.periodic[axis]) continue;
          p[axis] = (p[axis] + ext) % ext;
        }
        // a coarser leaf at level <= k.level-2 covering p must be split
        for (int up = 2; up <= k.level; up++) {
          Key c{k.level - up, p[0] >> up, p[1] >> up, p[2] >> up};
          if (set.count(c)) {
            split.push_back(c);
            break;
          }
        }
      }
    }
    std::sort(split.begin(), split.end(), [](const Key& a, const Key& b) {
      return std::tie(a.level, a.x, a.y, a.z) < std::tie(b.level, b.x, b.y, b.z);
    });
    split.erase(std::unique(split.begin(), split.end()), split.end());
    for (const Key& k : split) {
      if (!set.erase(k)) continue;
      for (int c = 0; c < 8; c++)
        set.insert({k.level + 1, 2 * k.x + (c & 1), 2 * k.y + ((c >> 1) & 1), 2 * k.z + ((c >> 2) & 1)});
      changed = true;
    }
  }
  blocks_.clear();
  max_level_ = 0;
  for (const Key& k : set) {
    Block b{};
    b.level = k.level;
    b.ix = k.x;
    b.iy = k.y;
    b.iz = k.z;
    blocks_.push_back(b);
    max_level_ = std::max(max_level_, k.level);
  }
}

void Octree::order() {
  const int L = max_level_;
  struct SortKey {
    int64_t root;
    uint64_t curve;
    size_t idx;
  };
  std::vector<SortKey> keys(blocks_.size());
  for (size_t i = 0; i < blocks_.size(); i++) {
    const Block& b = blocks_[i];
    int64_t rx = b.ix >> b.level, ry = b.iy >> b.level, rz = b.iz >> b.level;
    int64_t root = rx + spec_.root[0] * (ry + (int64_t)spec_.root[1] * rz);
    int sh = L - b.level;
    uint32_t lx = (uint32_t)((b.ix - (rx << b.level)) << sh);
    uint32_t ly = (uint32_t)((b.iy - (ry << b.level)) << sh);
    uint32_t lz = (uint32_t)((b.iz - (rz << b.level)) << sh);
    uint64_t c = spec_.ordering == 0 ? morton_index(lx, ly, lz, L) : hilbert_index(lx, ly, lz, L);
    keys[i] = {root, c, i};
  }
  std::sort(keys.begin(), keys.end(), [](const SortKey& a, const SortKey& b) {
    return a.root != b.root ? a.root < b.root : a.curve < b.curve;
  });
  std::vector<Block> sorted;
  sorted.reserve(blocks_.size());
  for (const SortKey& k : keys) sorted.push_back(blocks_[k.idx]);
  blocks_.swap(sorted);
}

void Octree::link() {
  std::unordered_map<Key, int, KeyHash> id;
  for (size_t i = 0; i < blocks_.size(); i++) id[{blocks_[i].level, blocks_[i].ix, blocks_[i].iy, blocks_[i].iz}] = (int)i;
  for (size_t i = 0; i < blocks_.size(); i++) {
    Block& b = blocks_[i];
    b.size = spec_.root_length / double(int64_t(1) << b.level);
    b.dx = b.size / spec_.n;
    b.org[0] = spec_.origin[0] + b.ix * b.size;
    b.org[1] = spec_.origin[1] + b.iy * b.size;
    b.org[2] = spec_.origin[2] + b.iz * b.size;
    for (int f = 0; f < 6; f++) {
      for (int s = 0; s < 4; s++) b.nbr[f][s] = -1;
      b.sub[f] = -1;
      b.ftype[f] = FACE_OUTER;
      int axis = f / 2;
      int64_t p[3] = {b.ix, b.iy, b.iz};
      p[axis] += (f & 1) ? 1 : -1;
      int64_t ext = (int64_t)spec_.root[axis] << b.level;
      if (p[axis] < 0 || p[axis] >= ext) {
        if (!spec_.periodic[axis]) continue;
        p[axis] = (p[axis] + ext) % ext;
      }
      int t0, t1;
      tangential(f, t0, t1);
      auto same = id.find({b.level, p[0], p[1], p[2]});
      if (same != id.end()) {
        b.ftype[f] = FACE_SAME;
        b.nbr[f][0] = same->second;
        continue;
      }
      if (b.level > 0) {
        auto coarse = id.find({b.level - 1, p[0] >> 1, p[1] >> 1, p[2] >> 1});
        if (coarse != id.end()) {
          int64_t me[3] = {b.ix, b.iy, b.iz};
          b.ftype[f] = FACE_COARSER;
          b.nbr[f][0] = coarse->second;
          b.sub[f] = (int)(me[t0] & 1) + 2 * (int)(me[t1] & 1);
          continue;
        }
      }
      b.ftype[f] = FACE_FINER;
      for (int s = 0; s < 4; s++) {
        int64_t c[3] = {2 * p[0], 2 * p[1], 2 * p[2]};
        c[axis] += (f & 1) ? 0 : 1;  // the children that touch this face
        c[t0] += s & 1;
        c[t1] += s >> 1;
        auto fine = id.find({b.level + 1, c[0], c[1], c[2]});
        if (fine == id.end()) throw std::runtime_error("octree is not 2:1 balanced");
        b.nbr[f][s] = fine->second;
      }
    }
  }
}

}  // namespace ffvhc
