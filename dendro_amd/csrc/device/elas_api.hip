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

// Host plan of the elasticity passes: elements sorted by stencil type (stable, so Morton
// order inside a type), cut into wavefront segments of one type; the node CSR of
// (sorted position, vertex) slots.
static void elas_plan(dgh_mesh m) {
  ElasState& S = m->elas;
  if (S.ready) return;
  const uint32_t E = m->E, N = m->N;
  std::vector<uint32_t> order(E);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return m->h_cs[a] < m->h_cs[b]; });
  std::vector<int4> seg;
  for (uint32_t i = 0; i < E;) {
    const uint8_t t = m->h_cs[order[i]];
    uint32_t j = i;
    while (j < E && j - i < uint32_t(ELAS_WAVE) && m->h_cs[order[j]] == t) ++j;
    seg.push_back(make_int4(int(t), int(i), int(j - i), 0));
    i = j;
  }
  std::vector<uint32_t> rowptr(size_t(N) + 1, 0);
  for (uint32_t q = 0; q < size_t(E) * 8; ++q) rowptr[m->h_nl[q] + 1]++;
  for (uint32_t v = 0; v < N; ++v) rowptr[v + 1] += rowptr[v];
  std::vector<uint32_t> slot(size_t(E) * 8), fill(rowptr.begin(), rowptr.end() - 1);
  for (uint32_t i = 0; i < E; ++i)
    for (int k = 0; k < 8; ++k) slot[fill[m->h_nl[size_t(order[i]) * 8 + k]]++] = i * 8 + uint32_t(k);
  if (E > (1u << 29)) fail(DGH_ERR_ARG, "too many elements for the elasticity slot encoding");
  S.order.upload(order.data(), E);
  S.seg.upload(seg.data(), seg.size());
  S.nseg = int(seg.size());
  S.rowptr.upload(rowptr.data(), rowptr.size());
  S.slot.upload(slot.data(), slot.size());
  S.lev.upload(m->h_lev.data(), E);
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
                         (const uint32_t*)S.order.p, (const uint32_t*)m->nl.p, (const uint8_t*)S.lev.p, m->maxD,
                         m->hFac, S.mu, S.lam, (const double*)g.EL2.p, (const double*)g.EGD.p, bdy, x, S.ye.p);
    if (m->N)
      hipLaunchKernelGGL(elas_node_kernel, dim3(blocks_for(m->N)), dim3(VBLOCK), 0, st, m->N,
                         (const uint32_t*)S.rowptr.p, (const uint32_t*)S.slot.p, (const double*)S.ye.p, bdy, x, y);
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
