// Contract of the owner-computes tile builder (dendro_amd/csrc/host/tiles.cpp), checked on
// the CPU for the budgets and owned-node capacities the operator kernels are launched with.
// The kernel's LDS layout relies on every one of these properties:
//   - n_own <= own_cap and own_cap + n_touch <= budget (own_cap = 0: n_own + n_touch <= budget);
//   - every node is owned by exactly one tile, listed first in its node table;
//   - a tile's records are every element touching one of its owned nodes (own or halo);
//   - record vertex ids index the tile's node table, in the vertex order k ^ child;
//   - records are sorted by stencil class;
//   - the device record form fits: vertex ids in 14 bits (a tile touches at most 16384 nodes)
//     and the class (child-0 type | level << 5) in the 10 bits the ids leave free.
// Usage: tile_plan_check npts depth   (prints one line per plan; exit status 1 on a failure)
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "dendro_hip.h"
#include "host/octree.hpp"

using namespace dgh;

static int failures = 0;
#define CHECK(c, ...)                  \
  do {                                 \
    if (!(c)) {                        \
      if (failures++ < 10) {           \
        std::printf("FAIL: " __VA_ARGS__); \
        std::printf("\n");             \
      }                                \
    }                                  \
  } while (0)

int main(int argc, char** argv) {
  const uint64_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 20000;
  const uint32_t D = argc > 2 ? uint32_t(std::atoi(argv[2])) : 9;
  std::vector<double> pts(3 * n);
  if (dgh_points_generate(1, n, 7, 0, pts.data()) != 0) return 2;
  const std::vector<Oct> tree = octree_from_points(pts.data(), n, D, 1, true);
  DALevel da;
  build_da(tree, D, da);
  const uint32_t E = da.n_elem, N = da.n_node;
  // elements touching each node
  std::vector<std::vector<uint32_t>> touching(N);
  for (uint32_t e = 0; e < E; ++e)
    for (int k = 0; k < 8; ++k) {
      auto& v = touching[da.nlist[size_t(e) * 8 + k]];
      if (v.empty() || v.back() != e) v.push_back(e);
    }
  struct Case { uint32_t T, budget, cap; };
  const Case cases[] = {{8192, 10176, 4096}, {4096, 10176, 4096}, {1024, 10176, 4096}, {2048, 5056, 2048},
                        {512, 2496, 1024}, {128, 2496, 1024}, {2048, 10176, 0}};
  for (const Case& c : cases) {
    TilePlan p;
    build_tiles(E, N, da.nlist.data(), da.child.data(), da.hnmask.data(), da.lev.data(), c.T, p, c.budget, c.cap);
    std::vector<int> owner_count(N, 0);
    std::vector<uint8_t> elem_seen(E, 0);
    size_t R = 0;
    for (size_t t = 0; t < p.tiles.size(); ++t) {
      const TileRec& tr = p.tiles[t];
      R += tr.n_elem;
      if (c.cap)
        CHECK(tr.n_own <= c.cap && c.cap + tr.n_touch <= c.budget, "tile %zu: own %u touch %u (cap %u budget %u)", t,
              tr.n_own, tr.n_touch, c.cap, c.budget);
      else
        CHECK(tr.n_own + tr.n_touch <= c.budget, "tile %zu: own %u + touch %u > budget %u", t, tr.n_own, tr.n_touch,
              c.budget);
      CHECK(tr.n_own <= tr.n_touch, "tile %zu: owned nodes not all touched", t);
      const uint32_t* nodes = p.node_tab.data() + tr.node_off;
      for (uint32_t i = 0; i < tr.n_own; ++i) owner_count[nodes[i]]++;
      std::set<uint32_t> recs;
      int prev_type = -1;
      for (uint32_t r = 0; r < tr.n_elem; ++r) {
        const size_t g = size_t(tr.elem_off) + r;
        const uint32_t e = p.rec_elem[g];
        recs.insert(e);
        elem_seen[e] = 1;
        const int type = p.meta[g] & 31;
        CHECK(type >= prev_type, "tile %zu: records not sorted by class", t);
        prev_type = type;
        CHECK((p.meta[g] >> 5) == da.lev[e], "record %zu: level", g);
        CHECK(p.meta[g] < 1024u, "record %zu: class %u beyond 10 bits", g, unsigned(p.meta[g]));
        for (int k = 0; k < 8; ++k) {
          const uint32_t l = p.lid[g * 8 + k];
          CHECK(l < tr.n_touch, "record %zu: vertex id %u beyond %u touched nodes", g, l, tr.n_touch);
          CHECK(l < 16384u, "record %zu: vertex id %u beyond 14 bits", g, l);
          if (l < tr.n_touch)
            CHECK(nodes[l] == da.nlist[size_t(e) * 8 + (k ^ da.child[e])], "record %zu vertex %d: node mismatch", g, k);
        }
      }
      for (uint32_t i = 0; i < tr.n_own; ++i)
        for (uint32_t e : touching[nodes[i]])
          CHECK(recs.count(e), "tile %zu: element %u touches owned node %u but has no record", t, e, nodes[i]);
    }
    for (uint32_t v = 0; v < N; ++v) CHECK(owner_count[v] == 1, "node %u owned %d times", v, owner_count[v]);
    for (uint32_t e = 0; e < E; ++e) CHECK(elem_seen[e], "element %u has no record", e);
    CHECK(R == p.rec_elem.size(), "record count");
    CHECK(p.max_touch <= 16384u, "max_touch %u beyond the 14-bit vertex ids", p.max_touch);
    std::printf("T=%u budget=%u cap=%u: E=%u N=%u tiles=%zu records=%zu max_own=%u max_touch=%u\n", c.T, c.budget,
                c.cap, E, N, p.tiles.size(), R, p.max_own, p.max_touch);
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
