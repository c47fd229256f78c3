// Linear octrees on the host: points -> octree, 2:1 balance, coarsening.
//
// Octants are (Morton index of the anchor at the tree's max depth D, level), kept sorted in
// Morton order.  The octree stages run once per mesh on the host, as in the reference
// (src/oct/Construct.C, src/oct/Balance.C, src/oct/Coarsen.C); only their results are
// uploaded to the device.
#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "octree.hpp"

namespace dgh {

// ---------------------------------------------------------------- points -> octree
// Driver preprocessing (examples/src/drivers/tstMg.C:148-192): quantise each point to its
// cell at depth D, keep it only if all coordinates are > 0 and the cell lies inside the
// cube, remove duplicate cells.  points2OctreeSeq (src/oct/Construct.C:884-930) then
// re-quantises the cell centres, which maps each centre back to its own cell.
static std::vector<uint64_t> point_cells(const double* xyz, uint64_t n, uint32_t D) {
  const double scale = double(1u << D);
  const uint32_t lim = 1u << D;
  std::vector<uint64_t> cells;
  cells.reserve(n);
  for (uint64_t i = 0; i < n; ++i) {
    const double px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    if (!(px > 0.0 && py > 0.0 && pz > 0.0)) continue;
    const uint32_t x = static_cast<uint32_t>(px * scale), y = static_cast<uint32_t>(py * scale),
                   z = static_cast<uint32_t>(pz * scale);
    if (x >= lim || y >= lim || z >= lim) continue;
    cells.push_back(morton(x, y, z));
  }
  std::sort(cells.begin(), cells.end());
  cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
  return cells;
}

// Coarsest complete linear octree whose leaves hold at most max_pts of the sorted point
// cells (recursive subdivision, depth <= D).  Same leaves as p2oLocal
// (src/oct/Construct.C:940-1010): walk octants in Morton order, descend through first
// children while an octant holds more than max_pts points, emit, move to the next octant
// at the coarsest level (TreeNode::getNext, include/oct/TreeNode.txx:435-528).
static std::vector<Oct> points_to_octree(const std::vector<uint64_t>& cells, uint32_t D,
                                         uint32_t max_pts) {
  std::vector<Oct> out;
  if (cells.size() <= max_pts) {
    out.push_back(Oct{0, 0, 0});
    return out;
  }
  struct Item { uint64_t m; uint8_t lev; size_t lo, hi; };
  std::vector<Item> stack;
  stack.push_back(Item{0, 0, 0, cells.size()});
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    if (it.hi - it.lo <= max_pts || it.lev == D) {
      out.push_back(Oct{it.m, it.lev, 0});
      continue;
    }
    const uint64_t sp = span(D, it.lev + 1);
    size_t bounds[9];
    bounds[0] = it.lo;
    for (int c = 1; c < 8; ++c)
      bounds[c] = std::lower_bound(cells.begin() + it.lo, cells.begin() + it.hi, it.m + c * sp) -
                  cells.begin();
    bounds[8] = it.hi;
    for (int c = 7; c >= 0; --c)
      stack.push_back(Item{it.m + c * sp, uint8_t(it.lev + 1), bounds[c], bounds[c + 1]});
  }
  return out;
}

// ---------------------------------------------------------------- linearise / complete
// Drops every octant that is an ancestor of (or equal to) its successor; input sorted.
static void linearise(std::vector<Oct>& v, uint32_t D) {
  std::vector<Oct> out;
  out.reserve(v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    if (i + 1 < v.size()) {
      const Oct& a = v[i];
      const Oct& b = v[i + 1];
      if (b.m >= a.m && b.m < a.m + span(D, a.lev) && b.lev >= a.lev) continue;
    }
    out.push_back(v[i]);
  }
  v.swap(out);
}

// Appends to out the coarsest octants tiling the Morton range [a, b) of finest cells.
static void fill_range(uint64_t a, uint64_t b, uint32_t D, std::vector<Oct>& out) {
  while (a < b) {
    uint32_t k = D;   // try the largest aligned octant first: span 8^k
    while (k > 0) {
      const uint64_t s = uint64_t(1) << (3 * k);
      if ((a & (s - 1)) == 0 && a + s <= b) break;
      --k;
    }
    out.push_back(Oct{a, uint8_t(D - k), 0});
    a += uint64_t(1) << (3 * k);
  }
}

// Fills the gaps of a sorted linear octree with the coarsest octants (a complete octree).
static std::vector<Oct> complete(const std::vector<Oct>& lin, uint32_t D) {
  std::vector<Oct> out;
  out.reserve(lin.size() * 2);
  uint64_t cur = 0;
  for (const Oct& o : lin) {
    if (o.m > cur) fill_range(cur, o.m, D, out);
    out.push_back(Oct{o.m, o.lev, 0});
    cur = o.m + span(D, o.lev);
  }
  const uint64_t end = uint64_t(1) << (3 * D);
  if (cur < end) fill_range(cur, end, D, out);
  return out;
}

// ---------------------------------------------------------------- 2:1 balance
// Minimal 2:1 balanced refinement: for every octant of level l (original or added) the
// same-level neighbours of its parent must exist at level >= l-1, so they are added as
// seeds at level l-1, finest level first; the seeds are then linearised and completed
// with the coarsest octants.  Face, edge and (inc_corner) corner neighbours.  The result
// is the unique coarsest balanced refinement, the tree ot::balanceOctree builds
// (src/oct/Balance.C:31; Sundar, Sampath & Biros, SIAM J. Sci. Comput. 30(5), 2008).
std::vector<Oct> balance(const std::vector<Oct>& in, uint32_t D, bool inc_corner) {
  std::vector<std::vector<uint64_t>> bins(D + 1);
  for (const Oct& o : in) bins[o.lev].push_back(o.m);
  const uint32_t lim = 1u << D;
  for (uint32_t l = D; l >= 2; --l) {
    auto& b = bins[l];
    // unique parents of this level's octants
    std::vector<uint64_t> parents;
    parents.reserve(b.size() / 4 + 1);
    const uint64_t pmask = ~(span(D, l - 1) - 1);
    for (uint64_t m : b) parents.push_back(m & pmask);
    std::sort(parents.begin(), parents.end());
    parents.erase(std::unique(parents.begin(), parents.end()), parents.end());
    const int32_t ps = static_cast<int32_t>(side(D, l - 1));
    auto& up = bins[l - 1];
    for (uint64_t pm : parents) {
      uint32_t px, py, pz;
      demorton(pm, px, py, pz);
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int nz = (dx != 0) + (dy != 0) + (dz != 0);
            if (nz == 0 || (nz == 3 && !inc_corner)) continue;
            const int64_t x = int64_t(px) + dx * ps, y = int64_t(py) + dy * ps,
                          z = int64_t(pz) + dz * ps;
            if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) continue;
            up.push_back(morton(uint32_t(x), uint32_t(y), uint32_t(z)));
          }
    }
    std::sort(up.begin(), up.end());
    up.erase(std::unique(up.begin(), up.end()), up.end());
  }
  std::vector<Oct> seeds;
  size_t total = 0;
  for (auto& b : bins) total += b.size();
  seeds.reserve(total);
  for (uint32_t l = 0; l <= D; ++l)
    for (uint64_t m : bins[l]) seeds.push_back(Oct{m, uint8_t(l), 0});
  bins.clear();
  std::sort(seeds.begin(), seeds.end());
  seeds.erase(std::unique(seeds.begin(), seeds.end()), seeds.end());
  linearise(seeds, D);
  return complete(seeds, D);
}

// ---------------------------------------------------------------- coarsening
// One pass of ot::coarsenOctree (src/oct/Coarsen.C:53-106): scan for consecutive child-0
// octants; when at least 8 octants separate them, the first 8 are a complete leaf family
// and are replaced by their parent; everything else is copied.
std::vector<Oct> coarsen(const std::vector<Oct>& in, uint32_t D) {
  if (in.size() < 8) return in;
  std::vector<Oct> out;
  out.reserve(in.size() / 2 + 8);
  const size_t n = in.size();
  size_t prev = 0;
  if (in[0].lev == 0 || child_number(in[0], D) != 0)
    fail(DGH_ERR_ARG, "coarsen: first octant is not a child 0");
  while (prev < n) {
    size_t cur = prev + 1;
    while (cur < n && (in[cur].lev == 0 || child_number(in[cur], D) != 0)) ++cur;
    if (cur >= prev + 8) {
      const Oct& c0 = in[prev];
      const uint8_t pl = uint8_t(c0.lev - 1);
      out.push_back(Oct{c0.m & ~(span(D, pl) - 1), pl, 0});
      for (size_t i = prev + 8; i < cur; ++i) out.push_back(in[i]);
    } else {
      for (size_t i = prev; i < cur; ++i) out.push_back(in[i]);
    }
    prev = cur;
  }
  return out;
}

std::vector<Oct> octree_from_points(const double* xyz, uint64_t n, uint32_t D, uint32_t max_pts,
                                    bool inc_corner) {
  if (D < 1 || D > 20) fail(DGH_ERR_ARG, "max_depth %u out of range [1,20]", D);
  if (max_pts < 1) fail(DGH_ERR_ARG, "max_pts must be >= 1");
  std::vector<uint64_t> cells = point_cells(xyz, n, D);
  std::vector<Oct> lin = points_to_octree(cells, D, max_pts);
  cells.clear();
  cells.shrink_to_fit();
  return balance(lin, D, inc_corner);
}

bool is_complete_linear(const std::vector<Oct>& v, uint32_t D) {
  uint64_t cur = 0;
  for (const Oct& o : v) {
    if (o.lev > D) return false;
    if (o.m != cur) return false;
    if (o.m & (span(D, o.lev) - 1)) return false;
    cur += span(D, o.lev);
  }
  return cur == (uint64_t(1) << (3 * D));
}

}  // namespace dgh
