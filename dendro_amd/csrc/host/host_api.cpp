// extern "C" host entry points of libdendrohip: stencil tables, synthetic points, octrees,
// and the multigrid hierarchy of DAs.
#include <cmath>
#include <cstdio>
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

// ---------------------------------------------------------------- hierarchy blob
// Versioned binary of a DA hierarchy: header {magic "DGHDAMG\0", version, nlevels, byte order
// mark}, then per level (coarsest first) the scalars {D, n_elem, n_node, n_hanging} and the
// arrays octs (Morton anchor u64, level u8, flags u8), node_of, nlist, child, hnmask, lev,
// and for levels >= 1 the transfer arrays celem, pos, first_fine, same; every array as
// {u64 count, count elements}; last a 64-bit FNV-1a hash of everything before it.
namespace {
constexpr char kBlobMagic[8] = {'D', 'G', 'H', 'D', 'A', 'M', 'G', 0};
constexpr uint32_t kBlobVersion = 1;
constexpr uint32_t kBlobBom = 0x01020304u;

struct BlobWriter {
  std::FILE* f;
  uint64_t h = 1469598103934665603ull;
  void raw(const void* p, size_t n) {
    const auto* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    if (n && std::fwrite(p, 1, n, f) != n) fail(DGH_ERR_ARG, "write failed");
  }
  template <class T> void val(T v) { raw(&v, sizeof(T)); }
  template <class T> void vec(const std::vector<T>& v) {
    val<uint64_t>(v.size());
    raw(v.data(), v.size() * sizeof(T));
  }
};

struct BlobReader {
  std::FILE* f;
  uint64_t h = 1469598103934665603ull;
  void raw(void* p, size_t n) {
    if (n && std::fread(p, 1, n, f) != n) fail(DGH_ERR_ARG, "hierarchy blob truncated");
    const auto* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  }
  template <class T> T val() {
    T v;
    raw(&v, sizeof(T));
    return v;
  }
  template <class T> void vec(std::vector<T>& v, uint64_t expect = ~0ull) {
    const uint64_t n = val<uint64_t>();
    if (n > (uint64_t(1) << 40) || (expect != ~0ull && n != expect)) fail(DGH_ERR_ARG, "hierarchy blob: bad array length");
    v.resize(n);
    raw(v.data(), n * sizeof(T));
  }
};

struct FileCloser {
  std::FILE* f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};
}  // namespace

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

int dgh_damg_save(dgh_damg h, const char* path) {
  return guarded([&] {
    if (!h || !path) fail(DGH_ERR_ARG, "NULL argument");
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) fail(DGH_ERR_ARG, "cannot open %s for writing", path);
    BlobWriter w{fc.f};
    w.raw(kBlobMagic, sizeof(kBlobMagic));
    w.val<uint32_t>(kBlobVersion);
    w.val<uint32_t>(uint32_t(h->levels.size()));
    w.val<uint32_t>(kBlobBom);
    for (size_t l = 0; l < h->levels.size(); ++l) {
      const DALevel& d = h->levels[l];
      w.val<uint32_t>(d.D);
      w.val<uint32_t>(d.n_elem);
      w.val<uint32_t>(d.n_node);
      w.val<uint32_t>(d.n_hanging);
      w.val<uint64_t>(d.octs.size());
      for (const Oct& o : d.octs) {
        w.val<uint64_t>(o.m);
        w.val<uint8_t>(o.lev);
        w.val<uint8_t>(o.flags);
      }
      w.vec(d.node_of);
      w.vec(d.nlist);
      w.vec(d.child);
      w.vec(d.hnmask);
      w.vec(d.lev);
      if (l > 0) {
        w.vec(h->celem[l]);
        w.vec(h->pos[l]);
        w.vec(h->first_fine[l]);
        w.vec(h->same[l]);
      }
    }
    const uint64_t hash = w.h;
    w.val<uint64_t>(hash);
    if (std::fflush(fc.f) != 0) fail(DGH_ERR_ARG, "write failed");
  });
}

int dgh_damg_load(const char* path, dgh_damg* out) {
  return guarded([&] {
    if (!path || !out) fail(DGH_ERR_ARG, "NULL argument");
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) fail(DGH_ERR_ARG, "cannot open %s", path);
    BlobReader r{fc.f};
    char magic[8];
    r.raw(magic, sizeof(magic));
    if (std::memcmp(magic, kBlobMagic, sizeof(magic)) != 0) fail(DGH_ERR_ARG, "%s is not a hierarchy blob", path);
    const uint32_t version = r.val<uint32_t>();
    if (version != kBlobVersion) fail(DGH_ERR_ARG, "hierarchy blob version %u (this library reads %u)", version, kBlobVersion);
    const uint32_t L = r.val<uint32_t>();
    if (r.val<uint32_t>() != kBlobBom) fail(DGH_ERR_ARG, "hierarchy blob of another byte order");
    if (L < 1 || L > 64) fail(DGH_ERR_ARG, "hierarchy blob: %u levels", L);
    auto h = std::make_unique<dgh_damg_s>();
    h->levels.resize(L);
    h->celem.resize(L);
    h->pos.resize(L);
    h->first_fine.resize(L);
    h->same.resize(L);
    for (uint32_t l = 0; l < L; ++l) {
      DALevel& d = h->levels[l];
      d.D = r.val<uint32_t>();
      d.n_elem = r.val<uint32_t>();
      d.n_node = r.val<uint32_t>();
      d.n_hanging = r.val<uint32_t>();
      if (d.D < 1 || d.D > 30) fail(DGH_ERR_ARG, "hierarchy blob: depth %u", d.D);
      const uint64_t no = r.val<uint64_t>();
      if (no < d.n_elem || no > (uint64_t(1) << 36)) fail(DGH_ERR_ARG, "hierarchy blob: bad octant count");
      d.octs.resize(no);
      for (Oct& o : d.octs) {
        o.m = r.val<uint64_t>();
        o.lev = r.val<uint8_t>();
        o.flags = r.val<uint8_t>();
      }
      r.vec(d.node_of, no);
      r.vec(d.nlist, uint64_t(d.n_elem) * 8);
      r.vec(d.child, d.n_elem);
      r.vec(d.hnmask, d.n_elem);
      r.vec(d.lev, d.n_elem);
      for (uint32_t v : d.nlist)
        if (v >= d.n_node) fail(DGH_ERR_MESH, "hierarchy blob: node id out of range");
      if (l > 0) {
        r.vec(h->celem[l], h->levels[l].n_node);
        r.vec(h->pos[l], h->levels[l].n_node);
        r.vec(h->first_fine[l], h->levels[l - 1].n_elem);
        r.vec(h->same[l], h->levels[l - 1].n_elem);
        for (uint32_t c : h->celem[l])
          if (c >= h->levels[l - 1].n_elem) fail(DGH_ERR_MESH, "hierarchy blob: coarse element out of range");
        for (uint8_t p : h->pos[l])
          if (p >= 27) fail(DGH_ERR_MESH, "hierarchy blob: position code out of range");
        for (uint32_t f : h->first_fine[l])
          if (f >= d.n_elem) fail(DGH_ERR_MESH, "hierarchy blob: fine element out of range");
      }
    }
    const uint64_t expect = r.h;
    if (r.val<uint64_t>() != expect) fail(DGH_ERR_ARG, "hierarchy blob checksum mismatch");
    *out = h.release();
  });
}

int dgh_damg_free(dgh_damg h) {
  delete h;
  return DGH_OK;
}

}  // extern "C"
