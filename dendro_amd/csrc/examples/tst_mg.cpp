// tst_mg: the flow of the reference's tstMg / tstMatVec drivers (examples/src/drivers/tstMg.C,
// examples/src/drivers/tstMatVec.C) through the C++ facade of libdendrohip: seeded points ->
// balanced octree -> DAMGCreateAndSetDA -> 100 timed Jacobian1 MatVecs on the finest level ->
// DAMGSetKSP (PCG + V-cycle, damped Jacobi) -> DAMGSolve.
//
//   tst_mg [npts=100000] [maxDepth=8] [nlevels=3] [kind=0] [--host-only]
//
// kind: 0 Gaussian (mean 0.5, sd 0.1), 1 log-normal, 2 uniform + clusters, 3 clusters,
// 4 uniform.  --host-only stops after the host hierarchy (no device needed).  The exit
// status is 0 when the solve meets a true relative residual of 1e-8.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dendro_hip.hpp"

using namespace dgh::omg;

// uniform [-1, 1) from splitmix64 (the right-hand side of the benchmark configurations is a
// seeded uniform vector)
static double uniform(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static int rhs_uniform(DAMG finest, double* b) {
  uint64_t s = 12;
  for (uint32_t i = 0; i < finest->info.n_node; ++i) b[i] = uniform(s);
  return 0;
}

#define CHECK(expr)                                                              \
  do {                                                                           \
    const int rc_ = (expr);                                                      \
    if (rc_ < 0) {                                                               \
      std::fprintf(stderr, "%s failed (%d): %s\n", #expr, rc_, lastError().c_str()); \
      return 2;                                                                  \
    }                                                                            \
  } while (0)

int main(int argc, char** argv) {
  uint64_t npts = 100000;
  uint32_t depth = 8;
  int nlevels = 3, kind = 0;
  bool host_only = false;
  int pos = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--host-only")) {
      host_only = true;
      continue;
    }
    const long v = std::atol(argv[i]);
    if (pos == 0) npts = uint64_t(v);
    else if (pos == 1) depth = uint32_t(v);
    else if (pos == 2) nlevels = int(v);
    else if (pos == 3) kind = int(v);
    ++pos;
  }
  std::vector<double> pts(3 * npts);
  CHECK(dgh_points_generate(kind, npts, 1, 31, pts.data()));
  dgh_octree oct = nullptr;
  CHECK(dgh_octree_from_points(pts.data(), npts, depth, 1, 1, &oct));
  uint64_t nleaf = 0;
  uint32_t maxd = 0;
  CHECK(dgh_octree_size(oct, &nleaf, &maxd));
  std::vector<uint32_t> xyz(3 * nleaf);
  std::vector<uint8_t> lev(nleaf);
  CHECK(dgh_octree_get(oct, xyz.data(), lev.data()));
  dgh_octree_free(oct);
  std::vector<TreeNode> leaves(nleaf);
  for (uint64_t i = 0; i < nleaf; ++i) leaves[i] = TreeNode{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], lev[i]};
  std::printf("tst_mg: %llu points, depth %u -> %llu balanced octants\n", (unsigned long long)npts, depth,
              (unsigned long long)nleaf);
  if (host_only) {
    dgh_octree t = nullptr;
    dgh_damg h = nullptr;
    CHECK(dgh_octree_from_leaves(xyz.data(), lev.data(), nleaf, maxd, &t));
    CHECK(dgh_damg_build(t, &nlevels, 500, &h));
    for (int l = 0; l < nlevels; ++l) {
      dgh_level_info li;
      CHECK(dgh_damg_level_info(h, l, &li));
      std::printf("tst_mg: level %d: %u elements, %u nodes\n", l, li.n_elem, li.n_node);
    }
    dgh_damg_free(h);
    dgh_octree_free(t);
    return 0;
  }
  CHECK(DAMG_Initialize(0));
  DAMG* damg = nullptr;
  CHECK(DAMGCreateAndSetDA(nlevels, nullptr, &damg, leaves, maxd));
  DAMG F = damg[nlevels - 1];
  // tstMatVec: 100 MatVecs of the finest level on a seeded vector
  const uint32_t N = F->info.n_node;
  std::vector<double> xh(N);
  uint64_t s = 11;
  for (uint32_t i = 0; i < N; ++i) xh[i] = uniform(s);
  double *xd = nullptr, *yd = nullptr;
  CHECK(dgh_alloc(sizeof(double) * N, reinterpret_cast<void**>(&xd)));
  CHECK(dgh_alloc(sizeof(double) * N, reinterpret_cast<void**>(&yd)));
  CHECK(dgh_copy_to_device(xd, xh.data(), sizeof(double) * N));
  CHECK(Jacobian1MatMult(F, xd, yd));
  CHECK(dgh_synchronize());
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < 100; ++i) CHECK(Jacobian1MatMult(F, xd, yd));
  CHECK(dgh_synchronize());
  const double mv_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / 100;
  CHECK(dgh_free(xd));
  CHECK(dgh_free(yd));
  dgh_solver_params p;
  dgh_solver_default_params(&p);               // bin/options values
  p.rtol = 1e-8;
  p.omega = 2.0 / 3.0;
  p.nu_pre = p.nu_post = 2;
  p.max_it = 200;
  p.norm_type = DGH_NORM_UNPRECONDITIONED;
  p.use_graph = 1;
  CHECK(DAMGSetKSP(damg, p, rhs_uniform));
  dgh_solve_stats st;
  CHECK(DAMGSolve(damg, &st));
  std::printf("tst_mg: levels=%d E=%u N=%u matvec_us=%.2f its=%d converged=%d rel_true=%.3e solve_ms=%.3f\n",
              nlevels, F->info.n_elem, N, mv_us, st.iterations, st.converged, st.rel_true, st.seconds * 1e3);
  CHECK(DAMGDestroy(damg));
  CHECK(DAMG_Finalize());
  return (st.converged && st.rel_true <= 1e-8) ? 0 : 1;
}
