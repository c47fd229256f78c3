// Linear elasticity operator (3 dofs per node) behind the C ABI of include/dendro_hip.h: the
// reference's ElasticityMatMult / ElasticityMatGetDiagonal (examples/src/backend/elasticityJac.C).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "elas_kernels.hpp"
#include "internal.hpp"

namespace dgh {

constexpr uint32_t ELAS_TILE = 2048;   // element cap of an elasticity tile (the node budget cuts first)

// Owner-computes tiling of the elasticity operator (the scalar operator's plan, with 24-byte
// node slots: a smaller owned-node capacity and node budget) and the records' meta with the
// child number.
static void elas_plan(dgh_mesh m) {
  ElasState& S = m->elas;
  if (S.ready) return;
  build_tiles(m->E, m->N, m->h_nl.data(), m->h_child.data(), m->h_hn.data(), m->h_lev.data(), ELAS_TILE, S.plan,
              ELAS_BUDGET, ELAS_OWN_CAP);
  if (elas_lds_bytes(S.plan.max_touch) > size_t(g.max_lds))
    fail(DGH_ERR_ARG, "elasticity tiles need %zu bytes of LDS (limit %d)", elas_lds_bytes(S.plan.max_touch), g.max_lds);
  const size_t R = S.plan.rec_elem.size();
  std::vector<uint16_t> meta(R);
  for (size_t r = 0; r < R; ++r) {   // child-0 hanging mask | child << 8 | level << 11
    const uint32_t t = S.plan.meta[r] & 31u, lev = S.plan.meta[r] >> 5;
    if (t >= uint32_t(NSTENCIL) || lev > 31u) fail(DGH_ERR_MESH, "record class out of range");
    meta[r] = uint16_t(kEtypeMask[0][t] | (uint32_t(m->h_child[S.plan.rec_elem[r]]) << 8) | (lev << 11));
  }
  S.tiles.upload(S.plan.tiles.data(), S.plan.tiles.size());
  S.lid.upload(S.plan.lid.data(), S.plan.lid.size());
  S.meta.upload(meta.data(), R);
  S.node_tab.upload(S.plan.node_tab.data(), S.plan.node_tab.size());
  S.ready = true;
}

// The unit element's matrix K_0 = mu L_0 (x) I + (mu + lambda) GD_0 (child 0, no hanging
// vertex; elasticityJac.C:171-206 applies fac = h/2 per element), and the per-tile
// fixed-point bounds from the element matrices fac (mu L (x) I + (mu + lambda) GD) of the
// (type, level) classes the tiling holds: texp[t].x = e with every row sum of |fac K_t| over
// the tile's records below 2^e, texp[t].y = log2 bound of the terms per owned node.
static void elas_matrices(dgh_mesh m) {
  ElasState& S = m->elas;
  std::vector<double> L(9216), G(size_t(144) * 576);
  generate_stencils(L.data(), nullptr, nullptr, nullptr);
  generate_graddiv(G.data());
  std::vector<double> k0(576), ke(576), rowmax(NCLASS, 0.0);
  for (int r = 0; r < 24; ++r)
    for (int q = 0; q < 24; ++q)
      k0[r * 24 + q] = S.mu * ((r % 3 == q % 3) ? L[(r / 3) * 8 + q / 3] : 0.0) + (S.mu + S.lam) * G[r * 24 + q];
  std::vector<uint8_t> seen(NCLASS, 0);
  for (uint16_t c : S.plan.meta) {
    if (seen[c]) continue;
    seen[c] = 1;
    const uint32_t t = c & 31u, lev = c >> 5;
    const double fac = m->hFac * double(1u << (m->maxD - lev)) / 2.0;
    const double f1 = S.mu * fac, f2 = (S.mu + S.lam) * fac;
    double* K = ke.data();
    for (int r = 0; r < 24; ++r) {
      double rs = 0.0;
      for (int q = 0; q < 24; ++q) {
        const double lap = (r % 3 == q % 3) ? L[size_t(t) * 64 + (r / 3) * 8 + q / 3] : 0.0;
        K[r * 24 + q] = f1 * lap + f2 * G[size_t(t) * 576 + r * 24 + q];
        rs += std::fabs(K[r * 24 + q]);
      }
      rowmax[c] = std::max(rowmax[c], rs);
    }
  }
  std::vector<int2> te(S.plan.tiles.size());
  std::vector<uint32_t> terms;
  for (size_t t = 0; t < S.plan.tiles.size(); ++t) {
    const TileRec& tr = S.plan.tiles[t];
    terms.assign(tr.n_own, 0u);
    double mx = 0.0;
    for (uint32_t r = tr.elem_off; r < tr.elem_off + tr.n_elem; ++r) {
      mx = std::max(mx, rowmax[S.plan.meta[r]]);
      for (int k = 0; k < 8; ++k) {
        const uint16_t l = S.plan.lid[size_t(r) * 8 + k];
        if (l < tr.n_own) terms[l]++;
      }
    }
    uint32_t vmax = 1;
    for (uint32_t c : terms) vmax = std::max(vmax, c);
    int v = 0;
    while ((1u << v) < vmax) ++v;
    if (!std::isfinite(mx)) fail(DGH_ERR_ARG, "non-finite elasticity element matrices");
    int e = -1022;
    if (mx > 0.0) std::frexp(mx, &e);
    te[t] = make_int2(std::max(e, -1022), v);
  }
  S.ke.upload(k0.data(), k0.size());
  S.texp.upload(te.data(), te.size());
}

}  // namespace dgh

extern "C" {

int dgh_mesh_set_elasticity(dgh_mesh m, double mu, double lambda) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (!(mu > 0.0) || !(mu + lambda > 0.0) || !std::isfinite(lambda))
      fail(DGH_ERR_ARG, "Lame parameters need mu > 0 and mu + lambda > 0");
    elas_plan(m);
    m->elas.mu = mu;
    m->elas.lam = lambda;
    elas_matrices(m);
  });
}

int dgh_matvec_elasticity(dgh_mesh m, const double* x, double* y, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !x || !y) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    if (!m->elas.ready) fail(DGH_ERR_STATE, "dgh_mesh_set_elasticity has not been called on this mesh");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const ElasState& S = m->elas;
    const int nt = int(S.plan.tiles.size());
    if (nt) {
      const unsigned grid = unsigned(NXCD * ((nt + NXCD - 1) / NXCD));
      const size_t lds = elas_lds_bytes(S.plan.max_touch);
      hipLaunchKernelGGL(elas_tile_kernel, dim3(grid), dim3(ELAS_BLOCK), lds, st, (const TileRec*)S.tiles.p, nt,
                         (const uint16_t*)S.lid.p, (const uint16_t*)S.meta.p, (const uint32_t*)S.node_tab.p,
                         (const int2*)S.texp.p, (const double*)S.ke.p, 0u, m->maxD, m->hFac, (const uint8_t*)m->bdy.p,
                         x, y);
    }
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
