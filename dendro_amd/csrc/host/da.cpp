// One level of the octree mesh (the reference's ot::DA on a single process) and the
// fine-node -> coarse-element transfer map between two levels.
#include <algorithm>
#include <thread>

#include "octree.hpp"

namespace dgh {

namespace {

// Runs f(begin, end) over [0, n) on up to 16 host threads.
template <class F>
void parallel_for(size_t n, F&& f) {
  unsigned nt = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 65536 || nt == 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& t : th) t.join();
}

// exact (anchor, level) lookup in the sorted octant array
int64_t find_exact(const std::vector<Oct>& v, uint64_t m, uint8_t lev) {
  auto it = std::lower_bound(v.begin(), v.end(), Oct{m, lev, 0});
  if (it != v.end() && it->m == m && it->lev == lev) return it - v.begin();
  return -1;
}

}  // namespace

int64_t DALevel::find_cell(uint32_t x, uint32_t y, uint32_t z, bool elements_only) const {
  const uint64_t key = morton(x, y, z);
  const auto end = elements_only ? octs.begin() + n_elem : octs.end();
  // last octant with m <= key (octants of a linear octree have distinct anchors)
  auto it = std::upper_bound(octs.begin(), end, Oct{key, 255, 0});
  if (it == octs.begin()) return -1;
  --it;
  if (key < it->m + span(D + 1, it->lev)) return it - octs.begin();
  return -1;
}

void build_da(const std::vector<Oct>& tree, uint32_t D, DALevel& da) {
  if (tree.empty()) fail(DGH_ERR_ARG, "build_da: empty octree");
  if (D + 1 > 21) fail(DGH_ERR_ARG, "build_da: max depth %u too large", D);
  const uint32_t Dda = D + 1;
  const uint32_t lim = 1u << D;   // the unit cube spans [0, 2^D] in DA coordinates
  da = DALevel();
  da.D = D;
  da.n_elem = static_cast<uint32_t>(tree.size());

  // 1. Embed one level deeper and add the positive-boundary pseudo-octants
  //    (addBoundaryNodesType1, src/oct/OctFunctions.C:636-700).
  std::vector<Oct> pseudo;
  da.octs.reserve(tree.size() + tree.size() / 8 + 8);
  for (const Oct& o : tree) {
    uint32_t x, y, z;
    demorton(o.m, x, y, z);
    const uint32_t sz = side(D, o.lev);
    const uint8_t L = uint8_t(o.lev + 1);
    da.octs.push_back(Oct{o.m, L, 0});
    const bool px = (x + sz == lim), py = (y + sz == lim), pz = (z + sz == lim);
    if (!(px || py || pz)) continue;
    if (px) pseudo.push_back(Oct{morton(lim, y, z), L, OCT_PSEUDO});
    if (py) pseudo.push_back(Oct{morton(x, lim, z), L, OCT_PSEUDO});
    if (pz) pseudo.push_back(Oct{morton(x, y, lim), L, OCT_PSEUDO});
    if (px && py) pseudo.push_back(Oct{morton(lim, lim, z), L, OCT_PSEUDO});
    if (px && pz) pseudo.push_back(Oct{morton(lim, y, lim), L, OCT_PSEUDO});
    if (py && pz) pseudo.push_back(Oct{morton(x, lim, lim), L, OCT_PSEUDO});
    if (px && py && pz) pseudo.push_back(Oct{morton(lim, lim, lim), L, OCT_PSEUDO});
  }
  std::sort(pseudo.begin(), pseudo.end());
  pseudo.erase(std::unique(pseudo.begin(), pseudo.end()), pseudo.end());
  // every pseudo-octant has a coordinate 2^D, so it follows all elements in Morton order
  da.octs.insert(da.octs.end(), pseudo.begin(), pseudo.end());
  pseudo.clear();
  const size_t n_oct = da.octs.size();

  // 2. Regular-node flags (flagNodesType3, src/oct/OctFunctions.C:1501-1900): the anchor of
  //    a child-1/2/4 octant lies on an edge of its parent, of a child-3/5/6 octant on a face;
  //    it hangs when a neighbouring octant of the parent's level across that edge or face is
  //    a leaf.  Children 0 and 7 and level-1 octants are always regular.
  std::vector<uint8_t> hanging(n_oct, 0);
  parallel_for(n_oct, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const Oct& o = da.octs[i];
      if (o.lev <= 1) continue;
      const uint32_t cn = child_number(o, Dda);
      if (cn == 0 || cn == 7) continue;
      uint32_t x, y, z;
      demorton(o.m, x, y, z);
      const uint32_t sz = side(Dda, o.lev);
      const uint32_t pmask = ~(2 * sz - 1);
      // directions to test per child number: bit0 -x, bit1 -y, bit2 -z (combined = edge)
      static const int kDirs[8][3] = {{0, 0, 0}, {2, 4, 6}, {1, 4, 5}, {4, 0, 0},
                                      {1, 2, 3}, {2, 0, 0}, {1, 0, 0}, {0, 0, 0}};
      bool hang = false;
      for (int q = 0; q < 3 && !hang; ++q) {
        const int d = kDirs[cn][q];
        if (!d) break;
        if (((d & 1) && !x) || ((d & 2) && !y) || ((d & 4) && !z)) continue;
        const uint32_t nx = (d & 1) ? x - sz : x, ny = (d & 2) ? y - sz : y,
                       nz = (d & 4) ? z - sz : z;
        if (find_exact(da.octs, morton(nx & pmask, ny & pmask, nz & pmask), uint8_t(o.lev - 1)) >= 0)
          hang = true;
      }
      hanging[i] = hang;
    }
  });
  da.node_of.assign(n_oct, kNoNode);
  uint32_t nn = 0;
  for (size_t i = 0; i < n_oct; ++i) {
    if (!hanging[i]) {
      da.octs[i].flags |= OCT_NODE;
      da.node_of[i] = nn++;
    }
  }
  da.n_node = nn;

  // 3. Element -> node lists.  Vertex k of an element is the anchor of the octant holding
  //    it when that octant's anchor is a regular node; otherwise the vertex hangs and is
  //    replaced by vertex k of the element's parent (odaBuildNlist.C:1835-1875,
  //    include/oda/oda.txx:1155-1190).
  const uint32_t E = da.n_elem;
  da.nlist.assign(size_t(E) * 8, 0);
  da.child.assign(E, 0);
  da.hnmask.assign(E, 0);
  da.lev.assign(E, 0);
  std::vector<int> err(1, 0);
  parallel_for(E, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      const Oct& o = da.octs[i];
      uint32_t x, y, z;
      demorton(o.m, x, y, z);
      const uint32_t sz = side(Dda, o.lev);
      const uint32_t px = x & ~(2 * sz - 1), py = y & ~(2 * sz - 1), pz = z & ~(2 * sz - 1);
      uint8_t mask = 0;
      for (int k = 0; k < 8; ++k) {
        const uint32_t qx = x + ((k & 1) ? sz : 0), qy = y + ((k & 2) ? sz : 0),
                       qz = z + ((k & 4) ? sz : 0);
        int64_t j = da.find_cell(qx, qy, qz, false);
        if (j >= 0 && da.octs[j].m == morton(qx, qy, qz) && (da.octs[j].flags & OCT_NODE)) {
          da.nlist[i * 8 + k] = da.node_of[j];
          continue;
        }
        mask |= uint8_t(1u << k);
        const uint32_t rx = px + ((k & 1) ? 2 * sz : 0), ry = py + ((k & 2) ? 2 * sz : 0),
                       rz = pz + ((k & 4) ? 2 * sz : 0);
        j = da.find_cell(rx, ry, rz, false);
        if (j < 0 || da.octs[j].m != morton(rx, ry, rz) || !(da.octs[j].flags & OCT_NODE)) {
          err[0] = 1;
          da.nlist[i * 8 + k] = 0;
          continue;
        }
        da.nlist[i * 8 + k] = da.node_of[j];
      }
      const int cn = int(child_number(o, Dda));
      da.child[i] = uint8_t(cn);
      da.hnmask[i] = mask;
      da.lev[i] = o.lev;
      if (etype_of(cn, mask) < 0) err[0] = 2;
    }
  });
  if (err[0] == 1) fail(DGH_ERR_MESH, "build_da: a hanging vertex has no regular parent vertex (octree not 2:1 balanced?)");
  if (err[0] == 2) fail(DGH_ERR_MESH, "build_da: hanging mask outside the 18 element types (octree not 2:1 balanced?)");
  uint32_t nh = 0;
  for (uint32_t i = 0; i < E; ++i) nh += da.hnmask[i] != 0;
  da.n_hanging = nh;
}

void match_levels(const DALevel& coarse, const DALevel& fine, std::vector<uint32_t>& first_fine,
                  std::vector<uint8_t>& same) {
  const uint32_t Ec = coarse.n_elem, Ef = fine.n_elem;
  first_fine.assign(Ec, 0);
  same.assign(Ec, 0);
  uint32_t f = 0;
  for (uint32_t c = 0; c < Ec; ++c) {
    const Oct& co = coarse.octs[c];
    if (f >= Ef || fine.octs[f].m != co.m)
      fail(DGH_ERR_MESH, "levels do not nest: coarse element %u has no matching fine element", c);
    first_fine[c] = f;
    if (fine.octs[f].lev == co.lev) {
      same[c] = 1;
      f += 1;
    } else if (fine.octs[f].lev == co.lev + 1) {
      const uint64_t sp = span(fine.D + 1, co.lev + 1);
      for (int k = 0; k < 8; ++k) {
        if (f + k >= Ef || fine.octs[f + k].m != co.m + k * sp || fine.octs[f + k].lev != co.lev + 1)
          fail(DGH_ERR_MESH, "levels do not nest: coarse element %u is not covered by 8 children", c);
      }
      f += 8;
    } else {
      fail(DGH_ERR_MESH, "levels do not nest: level jump at coarse element %u", c);
    }
  }
  if (f != Ef) fail(DGH_ERR_MESH, "levels do not nest: %u fine elements left over", Ef - f);
}

void build_transfer(const DALevel& coarse, const DALevel& fine, std::vector<uint32_t>& celem,
                    std::vector<uint8_t>& pos) {
  if (coarse.D != fine.D) fail(DGH_ERR_ARG, "transfer: levels of different max depth");
  const uint32_t lim = 1u << fine.D;
  const size_t n_oct = fine.octs.size();
  celem.assign(fine.n_node, 0);
  pos.assign(fine.n_node, 0);
  std::vector<int> err(1, 0);
  parallel_for(n_oct, [&](size_t b, size_t e) {
    for (size_t j = b; j < e; ++j) {
      const uint32_t node = fine.node_of[j];
      if (node == kNoNode) continue;
      uint32_t x, y, z;
      demorton(fine.octs[j].m, x, y, z);
      // a node on a positive face is found through the element on the inner side
      const uint32_t qx = (x == lim) ? x - 1 : x, qy = (y == lim) ? y - 1 : y,
                     qz = (z == lim) ? z - 1 : z;
      const int64_t c = coarse.find_cell(qx, qy, qz, true);
      if (c < 0) { err[0] = 1; continue; }
      uint32_t ax, ay, az;
      demorton(coarse.octs[c].m, ax, ay, az);
      const uint32_t s = side(coarse.D + 1, coarse.octs[c].lev), h = s / 2;
      auto code = [&](uint32_t p, uint32_t a) -> int {
        const uint32_t t = p - a;
        if (t == 0) return 0;
        if (h && t == h) return 1;
        if (t == s) return 2;
        return -1;
      };
      const int tx = code(x, ax), ty = code(y, ay), tz = code(z, az);
      if (tx < 0 || ty < 0 || tz < 0) { err[0] = 2; continue; }
      celem[node] = static_cast<uint32_t>(c);
      pos[node] = static_cast<uint8_t>(tx + 3 * ty + 9 * tz);
    }
  });
  if (err[0] == 1) fail(DGH_ERR_MESH, "transfer: a fine node lies outside the coarse mesh");
  if (err[0] == 2) fail(DGH_ERR_MESH, "transfer: a fine node is not on the 3x3x3 lattice of its coarse element");
}

}  // namespace dgh
