// Owner-computes tiling of one mesh level for the atomic-free operator kernel.
//
// Elements are cut into tiles of consecutive Morton-ordered elements.  Every node is owned
// by exactly one tile (the tile of the first element, in Morton order, that has the node
// as a vertex).  A tile's record list holds its own elements plus the "halo" elements of
// other tiles that touch a node it owns; the kernel evaluates those records, accumulates
// the rows of owned nodes in LDS and stores each owned node once.  Inside a tile the
// records are sorted by class (child-0 type, then level) so that the wavefronts read mostly
// one stencil.
//
// Stencils are mirror images across child numbers: reflecting the axes on which child c
// has its bit set maps the element to child 0 and vertex k to vertex k^c, so
// K[c][t][k][j] = K[0][t'][k^c][j^c] where t' is the child-0 type of the reflected hanging
// mask.  A record therefore stores its vertex list permuted by k -> k^c and the child-0
// type t', and the kernel needs the 18 child-0 stencils only.
#include <algorithm>
#include <numeric>
#include <thread>

#include "octree.hpp"

namespace dgh {

static uint8_t mirror_mask(uint8_t mask, int c) {
  uint8_t m = 0;
  for (int k = 0; k < 8; ++k)
    if (mask & (1u << k)) m |= uint8_t(1u << (k ^ c));
  return m;
}

// Greedy cut: a tile grows one element at a time.  Adding element b makes it an own
// element (its nodes become touched) and makes it the owner of the nodes whose first
// touching element it is; every later element touching such a node joins the tile's halo
// (its nodes become touched too).  Owned and touched counts only grow, so the tile ends
// before the first element whose addition would exceed the budget or the element cap.
// With own_cap > 0 the owned nodes have a fixed capacity: a tile also ends before it owns
// more than own_cap nodes or touches more than budget - own_cap.
void tile_bounds(uint32_t E, uint32_t N, const uint32_t* nl, uint32_t T, uint32_t budget,
                 std::vector<uint32_t>& bounds, uint32_t own_cap) {
  bounds.clear();
  if (budget == 0) {
    for (uint32_t e = 0; e < E; e += T) bounds.push_back(e);
    bounds.push_back(E);
    return;
  }
  std::vector<uint32_t> first(N, 0xffffffffu);
  for (uint32_t e = 0; e < E; ++e)
    for (int k = 0; k < 8; ++k) {
      uint32_t& f = first[nl[size_t(e) * 8 + k]];
      if (f == 0xffffffffu) f = e;
    }
  // elements touching each node (CSR), in ascending element order
  std::vector<uint32_t> ptr(size_t(N) + 1, 0), adj(size_t(E) * 8);
  for (size_t i = 0; i < size_t(E) * 8; ++i) ptr[nl[i] + 1]++;
  for (uint32_t n = 0; n < N; ++n) ptr[n + 1] += ptr[n];
  {
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    for (uint32_t e = 0; e < E; ++e)
      for (int k = 0; k < 8; ++k) adj[fill[nl[size_t(e) * 8 + k]]++] = e;
  }
  std::vector<uint32_t> nmark(N, 0), emark(E, 0);
  uint32_t stamp = 0;
  for (uint32_t a = 0; a < E;) {
    ++stamp;
    bounds.push_back(a);
    uint32_t own = 0, touch = 0, b = a;
    auto touch_elem = [&](uint32_t e) {
      if (emark[e] == stamp) return;
      emark[e] = stamp;
      for (int k = 0; k < 8; ++k) {
        const uint32_t n = nl[size_t(e) * 8 + k];
        if (nmark[n] != stamp) {
          nmark[n] = stamp;
          ++touch;
        }
      }
    };
    for (; b < E && b - a < T; ++b) {
      touch_elem(b);
      for (int k = 0; k < 8; ++k) {
        const uint32_t n = nl[size_t(b) * 8 + k];
        if (first[n] != b) continue;
        bool dup = false;   // a node listed twice in one element is owned once
        for (int q = 0; q < k; ++q) dup |= (nl[size_t(b) * 8 + q] == n);
        if (dup) continue;
        ++own;
        for (uint32_t p = ptr[n]; p < ptr[n + 1]; ++p)
          if (adj[p] > b) touch_elem(adj[p]);
      }
      if (touch > 65535u) break;
      if (own_cap ? (own > own_cap || own_cap + touch > budget) : (own + touch > budget)) break;
    }
    if (b == a) fail(DGH_ERR_ARG, "tile budget %u is smaller than one element's footprint", budget);
    a = b;
  }
  bounds.push_back(E);
}

void build_tiles(uint32_t E, uint32_t N, const uint32_t* nl, const uint8_t* child, const uint8_t* hn,
                 const uint8_t* lev, uint32_t T, TilePlan& plan, uint32_t budget, uint32_t own_cap) {
  if (T < 64 || T > 16384) fail(DGH_ERR_ARG, "tile size %u out of range [64,16384]", T);
  if (own_cap && own_cap >= budget) fail(DGH_ERR_ARG, "owned-node capacity %u not below the budget %u", own_cap, budget);
  plan = TilePlan();
  plan.tile_elems = T;
  plan.budget = budget;
  plan.own_cap = own_cap;
  std::vector<uint32_t> bounds;
  tile_bounds(E, N, nl, T, budget, bounds, own_cap);
  const uint32_t ntiles = uint32_t(bounds.size() - 1);
  std::vector<uint32_t> tile_of(E);
  for (uint32_t t = 0; t < ntiles; ++t)
    for (uint32_t e = bounds[t]; e < bounds[t + 1]; ++e) tile_of[e] = t;
  // owner tile of every node: first touching element's tile
  std::vector<uint32_t> owner(N, 0xffffffffu);
  for (uint32_t e = 0; e < E; ++e)
    for (int k = 0; k < 8; ++k) {
      uint32_t& o = owner[nl[size_t(e) * 8 + k]];
      if (o == 0xffffffffu) o = tile_of[e];
    }
  for (uint32_t n = 0; n < N; ++n)
    if (owner[n] == 0xffffffffu) fail(DGH_ERR_MESH, "node %u is not a vertex of any element", n);
  // halo elements per tile
  std::vector<std::vector<uint32_t>> halo(ntiles);
  for (uint32_t e = 0; e < E; ++e) {
    const uint32_t te = tile_of[e];
    uint32_t seen[8];
    int ns = 0;
    for (int k = 0; k < 8; ++k) {
      const uint32_t o = owner[nl[size_t(e) * 8 + k]];
      if (o == te) continue;
      bool dup = false;
      for (int q = 0; q < ns; ++q) dup |= (seen[q] == o);
      if (!dup) {
        seen[ns++] = o;
        halo[o].push_back(e);
      }
    }
  }
  // owned nodes per tile (ascending)
  std::vector<uint32_t> own_cnt(ntiles + 1, 0);
  for (uint32_t n = 0; n < N; ++n) own_cnt[owner[n] + 1]++;
  for (uint32_t t = 0; t < ntiles; ++t) own_cnt[t + 1] += own_cnt[t];
  std::vector<uint32_t> owned(N);
  {
    std::vector<uint32_t> fillp(own_cnt.begin(), own_cnt.end() - 1);
    for (uint32_t n = 0; n < N; ++n) owned[fillp[owner[n]]++] = n;
  }
  // element class t' (child-0 type of the mirrored mask)
  std::vector<uint8_t> tp(E);
  for (uint32_t e = 0; e < E; ++e) {
    const int t0 = etype_of(0, mirror_mask(hn[e], child[e]));
    if (t0 < 0) fail(DGH_ERR_MESH, "element %u: mirrored hanging mask is not a child-0 type", e);
    tp[e] = uint8_t(t0);
  }
  // per-tile records, built in parallel (each worker keeps a dense node -> local id map)
  plan.tiles.resize(ntiles);
  std::vector<std::vector<uint32_t>> t_rec(ntiles), t_nodes(ntiles);
  std::vector<std::vector<uint16_t>> t_lid(ntiles);
  unsigned nt = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (ntiles < 64) nt = 1;
  std::vector<std::thread> th;
  std::vector<int> err(nt, 0);
  for (unsigned w = 0; w < nt; ++w) {
    th.emplace_back([&, w] {
      std::vector<int32_t> lid_of(N, -1);
      for (uint32_t t = w; t < ntiles; t += nt) {
        auto& recs = t_rec[t];
        const uint32_t e0 = bounds[t], e1 = bounds[t + 1];
        recs.resize(e1 - e0);
        std::iota(recs.begin(), recs.end(), e0);
        recs.insert(recs.end(), halo[t].begin(), halo[t].end());
        // classes: child-0 type, then level (uniform coefficients read one element matrix
        // per (type, level) class)
        auto cls = [&](uint32_t e) { return (uint32_t(tp[e]) << 8) | lev[e]; };
        std::stable_sort(recs.begin(), recs.end(), [&](uint32_t a, uint32_t b) { return cls(a) < cls(b); });
        // Inside each class, deal the Morton-ordered records round-robin over the class's
        // wavefronts: the 64 lanes of a wavefront then hold elements about n/64 apart, so
        // siblings (which share the parent-centre node after the mirror permutation) do
        // not collide on one LDS address in the same atomic instruction.
        for (size_t a = 0; a < recs.size();) {
          size_t b = a;
          while (b < recs.size() && cls(recs[b]) == cls(recs[a])) ++b;
          const size_t n = b - a, W = (n + 63) / 64;
          if (W > 1) {
            std::vector<uint32_t> seg(recs.begin() + a, recs.begin() + b);
            size_t q = a;
            for (size_t w = 0; w < W; ++w)
              for (size_t i = w; i < n; i += W) recs[q++] = seg[i];
          }
          a = b;
        }
        auto& nodes = t_nodes[t];
        nodes.assign(owned.begin() + own_cnt[t], owned.begin() + own_cnt[t + 1]);
        for (size_t i = 0; i < nodes.size(); ++i) lid_of[nodes[i]] = int32_t(i);
        const size_t n_own = nodes.size();
        std::vector<uint32_t> extra;
        for (uint32_t e : recs)
          for (int k = 0; k < 8; ++k) {
            const uint32_t n = nl[size_t(e) * 8 + k];
            if (lid_of[n] == -1) {   // not owned and not yet listed
              lid_of[n] = -2;
              extra.push_back(n);
            }
          }
        std::sort(extra.begin(), extra.end());
        for (size_t i = 0; i < extra.size(); ++i) lid_of[extra[i]] = int32_t(n_own + i);
        nodes.insert(nodes.end(), extra.begin(), extra.end());
        if (nodes.size() > 65535) {
          err[w] = 1;
          break;
        }
        auto& lids = t_lid[t];
        lids.resize(recs.size() * 8);
        for (size_t r = 0; r < recs.size(); ++r) {
          const uint32_t e = recs[r];
          const int c = child[e];
          for (int k = 0; k < 8; ++k) lids[r * 8 + k] = uint16_t(lid_of[nl[size_t(e) * 8 + (k ^ c)]]);
        }
        for (uint32_t n : nodes) lid_of[n] = -1;
        plan.tiles[t].n_own = uint32_t(n_own);
        plan.tiles[t].n_touch = uint32_t(nodes.size());
        plan.tiles[t].n_elem = uint32_t(recs.size());
      }
    });
  }
  for (auto& x : th) x.join();
  for (int e : err)
    if (e) fail(DGH_ERR_ARG, "a tile touches more than 65535 nodes; use a smaller tile size");
  // concatenate
  size_t R = 0, NT = 0;
  for (uint32_t t = 0; t < ntiles; ++t) {
    plan.tiles[t].elem_off = uint32_t(R);
    plan.tiles[t].node_off = uint32_t(NT);
    R += t_rec[t].size();
    NT += t_nodes[t].size();
    plan.max_touch = std::max(plan.max_touch, plan.tiles[t].n_touch);
    plan.max_own = std::max(plan.max_own, plan.tiles[t].n_own);
    plan.max_elem = std::max(plan.max_elem, plan.tiles[t].n_elem);
    plan.max_slots = std::max(plan.max_slots, plan.tiles[t].n_own + plan.tiles[t].n_touch);
  }
  if (R > 0xffffffffu || NT > 0xffffffffu) fail(DGH_ERR_ARG, "mesh too large for 32-bit tile offsets");
  plan.rec_elem.resize(R);
  plan.meta.resize(R);
  plan.lid.resize(R * 8);
  plan.node_tab.resize(NT);
  for (uint32_t t = 0; t < ntiles; ++t) {
    const size_t ro = plan.tiles[t].elem_off, no = plan.tiles[t].node_off;
    std::copy(t_rec[t].begin(), t_rec[t].end(), plan.rec_elem.begin() + ro);
    std::copy(t_lid[t].begin(), t_lid[t].end(), plan.lid.begin() + ro * 8);
    std::copy(t_nodes[t].begin(), t_nodes[t].end(), plan.node_tab.begin() + no);
    for (size_t r = 0; r < t_rec[t].size(); ++r) {
      const uint32_t e = t_rec[t][r];
      plan.meta[ro + r] = uint16_t(tp[e] | (uint16_t(lev[e]) << 5));
    }
  }
}

}  // namespace dgh
