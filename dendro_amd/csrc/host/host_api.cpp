// extern "C" host entry points of libdendrohip: stencil tables, synthetic points, octrees,
// and the multigrid hierarchy of DAs.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "octree.hpp"

namespace dgh {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

// ---------------------------------------------------------------- random streams
// xoshiro256** (Blackman & Vigna 2018) seeded through splitmix64.
struct Rng {
  uint64_t s[4];
  bool have_spare = false;
  double spare = 0.0;
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9e3779b97f4a7c15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
  // Box-Muller: two uniforms give two independent N(0,1) variates
  double normal() {
    if (have_spare) {
      have_spare = false;
      return spare;
    }
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    spare = r * std::sin(th);
    have_spare = true;
    return r * std::cos(th);
  }
};

static void gen_points(int kind, uint64_t n, uint64_t seed, uint64_t cseed, double* xyz) {
  Rng rng(seed);
  double centres[8][3];
  if (kind == DGH_PTS_MIXED || kind == DGH_PTS_CLUSTERS) {
    Rng cr(cseed);
    for (auto& c : centres)
      for (double& v : c) v = 0.1 + 0.8 * cr.uniform();
  }
  switch (kind) {
    case DGH_PTS_GAUSS:
      for (uint64_t i = 0; i < 3 * n; ++i) xyz[i] = 0.5 + 0.1 * rng.normal();
      break;
    case DGH_PTS_UNIFORM:
      for (uint64_t i = 0; i < 3 * n; ++i) xyz[i] = rng.uniform();
      break;
    case DGH_PTS_LOGNORMAL: {
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
          const double v = std::exp(rng.normal());
          xyz[3 * i + d] = v;
          lo[d] = std::min(lo[d], v);
          hi[d] = std::max(hi[d], v);
        }
      for (uint64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d)
          xyz[3 * i + d] = (hi[d] > lo[d]) ? (xyz[3 * i + d] - lo[d]) / (hi[d] - lo[d]) : 0.5;
      break;
    }
    case DGH_PTS_MIXED:
    case DGH_PTS_CLUSTERS: {
      const uint64_t n_uni = (kind == DGH_PTS_MIXED) ? n / 2 : 0;
      for (uint64_t i = 0; i < 3 * n_uni; ++i) xyz[i] = rng.uniform();
      for (uint64_t i = n_uni; i < n; ++i) {
        const int c = int(rng.next() >> 61);
        for (int d = 0; d < 3; ++d) xyz[3 * i + d] = centres[c][d] + 0.02 * rng.normal();
      }
      break;
    }
    default:
      fail(DGH_ERR_ARG, "unknown point distribution %d", kind);
  }
}

}  // namespace dgh

using namespace dgh;

struct dgh_octree_s {
  std::vector<Oct> octs;
  uint32_t D = 0;
  bool inc_corner = true;
};

struct dgh_damg_s {
  std::vector<DALevel> levels;                    // coarsest first
  std::vector<std::vector<uint32_t>> celem;       // [level] fine-node -> coarse element
  std::vector<std::vector<uint8_t>> pos;
  std::vector<std::vector<uint32_t>> first_fine;  // [level] coarse element -> first fine element
  std::vector<std::vector<uint8_t>> same;
};

extern "C" {

int dgh_last_error(char* buf, size_t len) {
  if (!buf || !len) return DGH_ERR_ARG;
  std::strncpy(buf, g_last_error.c_str(), len - 1);
  buf[len - 1] = 0;
  return DGH_OK;
}

int dgh_etype(int child_num, int hn_mask) {
  const int t = etype_of(child_num, hn_mask);
  if (t < 0) {
    set_last_error("invalid (child number, hanging mask) pair");
    return DGH_ERR_MESH;
  }
  return t;
}

int dgh_stencils_generate(double* lap2, double* mass2, double* rmat2, double* rmat1) {
  return guarded([&] { generate_stencils(lap2, mass2, rmat2, rmat1); });
}

int dgh_points_generate(int kind, uint64_t n, uint64_t seed, uint64_t cluster_seed, double* xyz) {
  return guarded([&] {
    if (!xyz && n) fail(DGH_ERR_ARG, "xyz is NULL");
    gen_points(kind, n, seed, cluster_seed, xyz);
  });
}

int dgh_octree_from_points(const double* xyz, uint64_t n, uint32_t max_depth, uint32_t max_pts,
                           int inc_corner, dgh_octree* out) {
  return guarded([&] {
    if (!out || (!xyz && n)) fail(DGH_ERR_ARG, "NULL argument");
    auto t = std::make_unique<dgh_octree_s>();
    t->D = max_depth;
    t->inc_corner = inc_corner != 0;
    t->octs = octree_from_points(xyz, n, max_depth, max_pts, inc_corner != 0);
    *out = t.release();
  });
}

int dgh_octree_from_leaves(const uint32_t* xyz, const uint8_t* lev, uint64_t n, uint32_t max_depth,
                           dgh_octree* out) {
  return guarded([&] {
    if (!out || !xyz || !lev) fail(DGH_ERR_ARG, "NULL argument");
    if (max_depth < 1 || max_depth > 20) fail(DGH_ERR_ARG, "max_depth out of range");
    auto t = std::make_unique<dgh_octree_s>();
    t->D = max_depth;
    t->octs.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
      if (lev[i] > max_depth) fail(DGH_ERR_ARG, "leaf %llu: level above max_depth", (unsigned long long)i);
      t->octs[i] = Oct{morton(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), lev[i], 0};
    }
    if (!is_complete_linear(t->octs, max_depth))
      fail(DGH_ERR_ARG, "leaves are not a sorted complete linear octree");
    *out = t.release();
  });
}

int dgh_octree_balance(dgh_octree in, int inc_corner, dgh_octree* out) {
  return guarded([&] {
    if (!in || !out) fail(DGH_ERR_ARG, "NULL argument");
    auto t = std::make_unique<dgh_octree_s>();
    t->D = in->D;
    t->inc_corner = inc_corner != 0;
    t->octs = balance(in->octs, in->D, inc_corner != 0);
    *out = t.release();
  });
}

int dgh_octree_coarsen(dgh_octree in, dgh_octree* out) {
  return guarded([&] {
    if (!in || !out) fail(DGH_ERR_ARG, "NULL argument");
    auto t = std::make_unique<dgh_octree_s>();
    t->D = in->D;
    t->inc_corner = in->inc_corner;
    t->octs = coarsen(in->octs, in->D);
    *out = t.release();
  });
}

int dgh_octree_size(dgh_octree t, uint64_t* n, uint32_t* max_depth) {
  return guarded([&] {
    if (!t) fail(DGH_ERR_ARG, "NULL octree");
    if (n) *n = t->octs.size();
    if (max_depth) *max_depth = t->D;
  });
}

int dgh_octree_get(dgh_octree t, uint32_t* xyz, uint8_t* lev) {
  return guarded([&] {
    if (!t) fail(DGH_ERR_ARG, "NULL octree");
    for (size_t i = 0; i < t->octs.size(); ++i) {
      if (xyz) demorton(t->octs[i].m, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
      if (lev) lev[i] = t->octs[i].lev;
    }
  });
}

int dgh_octree_free(dgh_octree t) {
  delete t;
  return DGH_OK;
}

int dgh_damg_build(dgh_octree finest, int* nlevels, uint32_t min_coarse, dgh_damg* out) {
  return guarded([&] {
    if (!finest || !nlevels || !out) fail(DGH_ERR_ARG, "NULL argument");
    if (*nlevels < 1) fail(DGH_ERR_ARG, "nlevels must be >= 1");
    const uint32_t D = finest->D;
    // coarsen + balance until the requested count or a small level
    // (DAMGCreateAndSetDA, src/omg/omg.C:1363-1415)
    std::vector<std::vector<Oct>> trees;
    trees.push_back(finest->octs);
    while (int(trees.size()) < *nlevels) {
      std::vector<Oct> c = balance(coarsen(trees.back(), D), D, finest->inc_corner);
      const bool small = c.size() < min_coarse;
      trees.push_back(std::move(c));
      if (small) break;
    }
    auto h = std::make_unique<dgh_damg_s>();
    const int L = int(trees.size());
    h->levels.resize(L);
    for (int i = 0; i < L; ++i) {
      build_da(trees[L - 1 - i], D, h->levels[i]);   // level 0 = coarsest
      trees[L - 1 - i].clear();
      trees[L - 1 - i].shrink_to_fit();
    }
    h->celem.resize(L);
    h->pos.resize(L);
    h->first_fine.resize(L);
    h->same.resize(L);
    for (int i = 1; i < L; ++i) {
      build_transfer(h->levels[i - 1], h->levels[i], h->celem[i], h->pos[i]);
      match_levels(h->levels[i - 1], h->levels[i], h->first_fine[i], h->same[i]);
    }
    *nlevels = L;
    *out = h.release();
  });
}

int dgh_damg_num_levels(dgh_damg h, int* nlevels) {
  return guarded([&] {
    if (!h || !nlevels) fail(DGH_ERR_ARG, "NULL argument");
    *nlevels = int(h->levels.size());
  });
}

static const DALevel& level_of(dgh_damg h, int level) {
  if (!h) fail(DGH_ERR_ARG, "NULL hierarchy");
  if (level < 0 || level >= int(h->levels.size())) fail(DGH_ERR_ARG, "level %d out of range", level);
  return h->levels[level];
}

int dgh_damg_level_info(dgh_damg h, int level, dgh_level_info* info) {
  return guarded([&] {
    const DALevel& da = level_of(h, level);
    if (!info) fail(DGH_ERR_ARG, "NULL info");
    info->n_elem = da.n_elem;
    info->n_node = da.n_node;
    info->n_oct = uint32_t(da.octs.size());
    info->max_depth = da.max_depth();
    info->n_hanging = da.n_hanging;
  });
}

int dgh_damg_level_arrays(dgh_damg h, int level, uint32_t* nlist, uint8_t* child, uint8_t* hn_mask,
                          uint8_t* lev, uint32_t* anchor) {
  return guarded([&] {
    const DALevel& da = level_of(h, level);
    const size_t E = da.n_elem;
    if (nlist) std::memcpy(nlist, da.nlist.data(), E * 8 * sizeof(uint32_t));
    if (child) std::memcpy(child, da.child.data(), E);
    if (hn_mask) std::memcpy(hn_mask, da.hnmask.data(), E);
    if (lev) std::memcpy(lev, da.lev.data(), E);
    if (anchor)
      for (size_t i = 0; i < E; ++i)
        demorton(da.octs[i].m, anchor[3 * i], anchor[3 * i + 1], anchor[3 * i + 2]);
  });
}

int dgh_damg_level_octants(dgh_damg h, int level, uint32_t* anchor, uint8_t* lev, uint8_t* flags) {
  return guarded([&] {
    const DALevel& da = level_of(h, level);
    for (size_t i = 0; i < da.octs.size(); ++i) {
      if (anchor) demorton(da.octs[i].m, anchor[3 * i], anchor[3 * i + 1], anchor[3 * i + 2]);
      if (lev) lev[i] = da.octs[i].lev;
      if (flags) flags[i] = da.octs[i].flags;
    }
  });
}

// A node is on the domain boundary when its anchor has a coordinate 0 or 2^D (the unit cube
// spans [0, 2^D] in DA coordinates); every node is the anchor of exactly one octant.
int dgh_damg_level_boundary(dgh_damg h, int level, uint8_t* bdy) {
  return guarded([&] {
    if (!bdy) fail(DGH_ERR_ARG, "NULL argument");
    const DALevel& da = level_of(h, level);
    const uint32_t lim = 1u << da.D;
    for (size_t i = 0; i < da.octs.size(); ++i) {
      const uint32_t n = da.node_of[i];
      if (n == kNoNode) continue;
      uint32_t x, y, z;
      demorton(da.octs[i].m, x, y, z);
      bdy[n] = (x == 0 || y == 0 || z == 0 || x == lim || y == lim || z == lim) ? 1 : 0;
    }
  });
}

int dgh_damg_transfer_arrays(dgh_damg h, int fine_level, uint32_t* coarse_elem, uint8_t* pos) {
  return guarded([&] {
    const DALevel& da = level_of(h, fine_level);
    if (fine_level < 1) fail(DGH_ERR_ARG, "level 0 has no coarser level");
    if (coarse_elem) std::memcpy(coarse_elem, h->celem[fine_level].data(), size_t(da.n_node) * 4);
    if (pos) std::memcpy(pos, h->pos[fine_level].data(), da.n_node);
  });
}

int dgh_damg_coarsen_coefficients(dgh_damg h, const double* fine_kappa, const double* fine_mass,
                                  int level, double* out_kappa, double* out_mass) {
  return guarded([&] {
    if (!h) fail(DGH_ERR_ARG, "NULL hierarchy");
    const int L = int(h->levels.size());
    level_of(h, level);
    // walk from the finest level down to `level`
    std::vector<double> k(fine_kappa ? fine_kappa : nullptr,
                          fine_kappa ? fine_kappa + h->levels[L - 1].n_elem : nullptr);
    std::vector<double> m(fine_mass ? fine_mass : nullptr,
                          fine_mass ? fine_mass + h->levels[L - 1].n_elem : nullptr);
    for (int f = L - 1; f > level; --f) {
      const DALevel& c = h->levels[f - 1];
      std::vector<double> kc(k.empty() ? 0 : c.n_elem), mc(m.empty() ? 0 : c.n_elem);
      for (uint32_t e = 0; e < c.n_elem; ++e) {
        const uint32_t ff = h->first_fine[f][e];
        if (h->same[f][e]) {
          if (!k.empty()) kc[e] = k[ff];
          if (!m.empty()) mc[e] = m[ff];
        } else {
          double sk = 0.0, sm = 0.0;
          for (int q = 0; q < 8; ++q) {
            if (!k.empty()) sk += k[ff + q];
            if (!m.empty()) sm += m[ff + q];
          }
          if (!k.empty()) kc[e] = sk / 8.0;
          if (!m.empty()) mc[e] = sm / 8.0;
        }
      }
      k.swap(kc);
      m.swap(mc);
    }
    if (out_kappa && !k.empty()) std::memcpy(out_kappa, k.data(), k.size() * sizeof(double));
    if (out_mass && !m.empty()) std::memcpy(out_mass, m.data(), m.size() * sizeof(double));
  });
}

int dgh_damg_free(dgh_damg h) {
  delete h;
  return DGH_OK;
}

}  // extern "C"
