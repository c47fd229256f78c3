// Linear elasticity operator (3 dofs per node) behind the C ABI of include/dendro_hip.h: the
// reference's ElasticityMatMult / ElasticityMatGetDiagonal (examples/src/backend/elasticityJac.C).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "elas_kernels.hpp"
#include "internal.hpp"

namespace dgh {

static void elas_tables() {
  if (g.EL2.p) return;
  std::vector<double> L(9216), G(size_t(144) * 576);
  generate_stencils(L.data(), nullptr, nullptr, nullptr);
  generate_graddiv(G.data());
  g.EL2.upload(L.data(), L.size());
  g.EGD.upload(G.data(), G.size());
}

// Host plan of the elasticity passes: inside chunks of ELAS_CHUNK Morton-consecutive elements,
// elements sorted by their mirrored (child-0) stencil type, cut into wavefront segments of one
// type; the node CSR of (sorted position, vertex) slots.
static uint8_t mirrored_type(uint8_t hn, uint8_t c) {
  uint32_t mm = 0;
  for (int k = 0; k < 8; ++k)
    if (hn & (1u << k)) mm |= 1u << (k ^ c);
  const int t = etype_of(0, int(mm));
  if (t < 0) fail(DGH_ERR_MESH, "mirrored hanging mask is not a child-0 type");
  return uint8_t(t);
}

static void elas_plan(dgh_mesh m) {
  ElasState& S = m->elas;
  if (S.ready) return;
  const uint32_t E = m->E, N = m->N;
  if (E > (1u << 29)) fail(DGH_ERR_ARG, "too many elements for the elasticity slot encoding");
  std::vector<uint8_t> tp(E);
  for (uint32_t e = 0; e < E; ++e) tp[e] = mirrored_type(m->h_hn[e], m->h_child[e]);
  std::vector<uint32_t> order(E);
  std::iota(order.begin(), order.end(), 0u);
  std::vector<int4> seg;
  for (uint32_t c0 = 0; c0 < E; c0 += ELAS_CHUNK) {
    const uint32_t c1 = std::min<uint32_t>(E, c0 + ELAS_CHUNK);
    std::stable_sort(order.begin() + c0, order.begin() + c1, [&](uint32_t a, uint32_t b) { return tp[a] < tp[b]; });
    for (uint32_t i = c0; i < c1;) {
      const uint8_t t = tp[order[i]];
      uint32_t j = i;
      while (j < c1 && j - i < uint32_t(ELAS_WAVE) && tp[order[j]] == t) ++j;
      seg.push_back(make_int4(int(t), int(i), int(j - i), 0));
      i = j;
    }
  }
  std::vector<uint32_t> rowptr(size_t(N) + 1, 0);
  for (size_t q = 0; q < size_t(E) * 8; ++q) rowptr[m->h_nl[q] + 1]++;
  for (uint32_t v = 0; v < N; ++v) rowptr[v + 1] += rowptr[v];
  std::vector<uint32_t> slot(size_t(E) * 8), fill(rowptr.begin(), rowptr.end() - 1);
  for (uint32_t i = 0; i < E; ++i)
    for (int k = 0; k < 8; ++k) slot[fill[m->h_nl[size_t(order[i]) * 8 + k]]++] = i * 8 + uint32_t(k);
  S.order.upload(order.data(), E);
  S.seg.upload(seg.data(), seg.size());
  S.nseg = int(seg.size());
  S.rowptr.upload(rowptr.data(), rowptr.size());
  S.slot.upload(slot.data(), slot.size());
  S.lev.upload(m->h_lev.data(), E);
  S.child.upload(m->h_child.data(), E);
  S.ye.alloc(size_t(E) * 24);
  S.ready = true;
}

}  // namespace dgh

extern "C" {

int dgh_mesh_set_elasticity(dgh_mesh m, double mu, double lambda) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (!(mu > 0.0) || !(mu + lambda > 0.0) || !std::isfinite(lambda))
      fail(DGH_ERR_ARG, "Lame parameters need mu > 0 and mu + lambda > 0");
    elas_tables();
    elas_plan(m);
    m->elas.mu = mu;
    m->elas.lam = lambda;
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
    const uint8_t* bdy = m->bdy.p;
    if (S.nseg)
      hipLaunchKernelGGL(elas_elem_kernel, dim3((S.nseg + 3) / 4), dim3(256), 0, st, S.nseg, (const int4*)S.seg.p,
                         (const uint32_t*)S.order.p, (const uint32_t*)m->nl.p, (const uint8_t*)S.lev.p,
                         (const uint8_t*)S.child.p, m->maxD, m->hFac, S.mu, S.lam, (const double*)g.EL2.p,
                         (const double*)g.EGD.p, bdy, x, S.ye.p);
    if (m->N)
      hipLaunchKernelGGL(elas_node_kernel, dim3(blocks_for(m->N)), dim3(VBLOCK), 0, st, m->N,
                         (const uint32_t*)S.rowptr.p, (const uint32_t*)S.slot.p, (const double*)S.ye.p, bdy, x, y);
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
