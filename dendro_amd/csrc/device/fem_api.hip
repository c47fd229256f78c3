// Right-hand sides, nodal values and L2 errors of functions on the unit cube, behind the C ABI
// of include/dendro_hip.h (the reference's ComputeRHS2 / ComputeSol / ComputeError5,
// examples/src/backend/omgRhs.C).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "internal.hpp"
#include "rhs_kernels.hpp"

namespace dgh {

static FnDev fn_dev(const dgh_function* f) {
  if (!f) fail(DGH_ERR_ARG, "NULL function");
  if (f->kind != DGH_FN_GAUSSIANS && f->kind != DGH_FN_COS_SUM && f->kind != DGH_FN_POLY)
    fail(DGH_ERR_ARG, "unknown function kind %d", f->kind);
  if (f->kind == DGH_FN_COS_SUM && (f->nfreq < 0 || f->nfreq > 64)) fail(DGH_ERR_ARG, "nfreq out of range");
  FnDev d{f->kind, f->nfreq, f->a0, f->a1, {}};
  for (int i = 0; i < 8; ++i) d.c[i] = f->poly[i];
  return d;
}

static void require_geometry(dgh_mesh m) {
  if (m->E && m->geo.n != m->E) fail(DGH_ERR_STATE, "dgh_mesh_set_geometry has not been called on this mesh");
}

static const double* shfn_table() {
  if (!g.SF.p) {
    std::vector<double> S(size_t(144) * 8 * NQ);
    generate_shfn(S.data());
    g.SF.upload(S.data(), S.size());
  }
  return g.SF.p;
}

}  // namespace dgh

extern "C" {

int dgh_mesh_set_geometry(dgh_mesh m, const uint32_t* anchor) {
  return guarded([&] {
    require_init();
    if (!m || (m->E && !anchor)) fail(DGH_ERR_ARG, "NULL argument");
    const uint32_t lim = 1u << m->maxD;
    std::vector<int4> geo(m->E);
    for (uint32_t e = 0; e < m->E; ++e) {
      const uint32_t* a = anchor + size_t(e) * 3;
      if (a[0] >= lim || a[1] >= lim || a[2] >= lim) fail(DGH_ERR_MESH, "element %u: anchor outside the domain", e);
      geo[e] = make_int4(int(a[0]), int(a[1]), int(a[2]), int(m->h_lev[e]));
    }
    m->geo.upload(geo.data(), geo.size());
  });
}

int dgh_rhs_quadrature(dgh_mesh m, const dgh_function* f, double* b, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !b) fail(DGH_ERR_ARG, "NULL argument");
    const FnDev fd = fn_dev(f);
    require_geometry(m);
    const double* SF = shfn_table();
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m->N) hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks_for(m->N)), dim3(VBLOCK), 0, st, b, 0.0, m->N);
    if (m->E)
      hipLaunchKernelGGL(rhs_quad_kernel, dim3(blocks_for(m->E)), dim3(VBLOCK), 0, st, m->E, m->maxD,
                         (const int4*)m->geo.p, (const uint32_t*)m->nl.p, (const uint8_t*)m->cs.p, SF, fd, b);
    HIP_CHECK(hipGetLastError());
  });
}

int dgh_nodal_values(dgh_mesh m, const dgh_function* f, double* u, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !u) fail(DGH_ERR_ARG, "NULL argument");
    const FnDev fd = fn_dev(f);
    require_geometry(m);
    if (m->hn.n != m->E) m->hn.upload(m->h_hn.data(), m->E);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (m->E)
      hipLaunchKernelGGL(nodal_values_kernel, dim3(blocks_for(m->E)), dim3(VBLOCK), 0, st, m->E, m->maxD,
                         (const int4*)m->geo.p, (const uint32_t*)m->nl.p, (const uint8_t*)m->hn.p, fd, u);
    HIP_CHECK(hipGetLastError());
  });
}

int dgh_l2_error(dgh_mesh m, const dgh_function* f, const double* u, double* err) {
  return guarded([&] {
    require_init();
    if (!m || !u || !err) fail(DGH_ERR_ARG, "NULL argument");
    const FnDev fd = fn_dev(f);
    require_geometry(m);
    const double* SF = shfn_table();
    DevBuf<double> part(RED_BLOCKS + 1);
    hipLaunchKernelGGL(l2_error_kernel, dim3(RED_BLOCKS), dim3(VBLOCK), 0, 0, m->E, m->maxD, (const int4*)m->geo.p,
                       (const uint32_t*)m->nl.p, (const uint8_t*)m->cs.p, SF, fd, u, part.p);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(RED2_THREADS), 0, 0, int(RED_BLOCKS),
                       (const double*)part.p, part.p + RED_BLOCKS);
    HIP_CHECK(hipGetLastError());
    double s = 0.0;
    HIP_CHECK(hipMemcpy(&s, part.p + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost));
    *err = std::sqrt(s);
  });
}

}  // extern "C"
