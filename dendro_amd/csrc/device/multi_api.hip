// Batched right-hand sides behind the C ABI: the k-column operator, its tile autotuner and
// a CG over k independent columns preconditioned by one batched FP32 V-cycle per iteration
// (every column keeps its own CG scalars; the iteration is captured as one hipGraph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "internal.hpp"
#include "kernels_multi.hpp"

namespace dgh {

template <class T> static const T* dk(const DevTiles& d);
template <class T> static const T* dm(const DevTiles& d);
template <> const double* dk<double>(const DevTiles& d) { return d.rk64.p; }
template <> const double* dm<double>(const DevTiles& d) { return d.rm64.p; }
template <> const float* dk<float>(const DevTiles& d) { return d.rk32.p; }
template <> const float* dm<float>(const DevTiles& d) { return d.rm32.p; }

// LDS bytes of one batched workgroup: fixed-point sums of the owned nodes (vector-ALU kernel;
// FP64 sums in the matrix-core variant) and the staged x of the touched nodes, rows of a
// slice padded by one value (vector-ALU kernel: FP32 slices of 8 columns, FP64 slices of 4);
// the matrix-core variant adds the full stencil table.
template <class T, int KSL>
static size_t multi_lds_valu(const DevTiles& d) {
  return size_t(multi_row<T>(KSL)) * (sizeof(fx_acc_t<T>) * d.ys_len + sizeof(T) * d.plan.max_touch);
}
static size_t multi_lds_mfma(const DevTiles& d, bool mass) {
  return sizeof(double) * KS * d.ys_len + sizeof(float) * KS * d.plan.max_touch +
         sizeof(float) * multi_stencil_len<true>() / (mass ? 1 : 2);
}
static size_t multi_lds_worst(const DevTiles& d) {
  return std::max({multi_lds_valu<float, KS>(d), multi_lds_valu<double, KS64>(d), multi_lds_mfma(d, true)});
}

// FP32 product variants: 0 vector ALUs (8 columns per slice, 512 threads, one column per
// lane), 1 matrix cores, 2 as 0 with 1024 threads, 3 as 0 with 4 columns per slice (half the
// LDS), 4 and 5 as 0 with 2 and 4 columns per lane, 6 as 3 with 2 columns per lane
constexpr int NVARIANT = 7;

static void plan_multi(dgh_mesh m, uint32_t T) {
  build_devtiles(m, m->mop, T, multi_lds_worst);
  m->mtile = m->mop.plan.tile_elems;
}

void ensure_multi(dgh_mesh m) {
  if (m->mop.plan.tiles.empty()) plan_multi(m, m->mtile ? m->mtile : 1024);
}

static void check_k(int k) {
  if (k < KS || k > MAXK || k % KS) fail(DGH_ERR_ARG, "k = %d right-hand sides: must be a multiple of %d in [%d, %d]", k, KS, KS, MAXK);
}

template <class T, bool MASS, int MODE, int KSL, int MB, int CPL = 1>
static void launch_valu(dgh_mesh m, const DevTiles& d, const T* kap, const T* mas, const MultiIO<T>& io,
                        hipStream_t st) {
  const int nt = int(d.plan.tiles.size());
  const int np = nt * (io.k / KSL);
  const unsigned grid = unsigned(NXCD * ((np + NXCD - 1) / NXCD));
  const size_t lds = multi_lds_valu<T, KSL>(d);
  hipLaunchKernelGGL((tile_multi_kernel<T, MASS, MODE, KSL, MB, CPL>), dim3(grid), dim3(MB), lds, st,
                     d.tiles.p, nt, io.k / KSL, d.lid.p, kap, mas, m->lap_fac, m->mass_fac, m->maxD,
                     m->hFac, Ktab<T>(), d.node_tab.p, (const int2*)d.texp.p, d.ys_len, io);
}

template <bool MASS, int MODE>
static void launch_mfma(dgh_mesh m, const DevTiles& d, const float* kap, const float* mas,
                        const MultiIO<float>& io, hipStream_t st) {
  const int nt = int(d.plan.tiles.size());
  const int np = nt * (io.k / KS);
  const unsigned grid = unsigned(NXCD * ((np + NXCD - 1) / NXCD));
  hipLaunchKernelGGL((tile_multi_mfma_kernel<MASS, MODE, MBLOCK>), dim3(grid), dim3(MBLOCK), multi_lds_mfma(d, MASS), st,
                     d.tiles.p, nt, io.k / KS, d.lid.p, kap, mas, m->lap_fac, m->mass_fac, m->maxD,
                     m->hFac, Ktab<float>(), d.node_tab.p, d.ys_len, io);
}

// The batched kernels move 16-byte vectors of a node's row.
static void check_aligned(const void* a, const void* b = nullptr, const void* c = nullptr) {
  if ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(c)) & 15u)
    fail(DGH_ERR_ARG, "batched vectors must be 16-byte aligned");
}

template <class T, int MODE>
static void launch_multi(dgh_mesh m, const MultiIO<T>& io, hipStream_t st) {
  check_aligned(io.x, io.y, io.b);
  ensure_multi(m);
  const DevTiles& d = m->mop;
  if (d.plan.tiles.empty()) return;
  const T* kap = dk<T>(d);
  const T* mas = dm<T>(d);
  const bool mass = !(kap && !mas);
  if constexpr (std::is_same<T, double>::value) {
    if (mass) launch_valu<double, true, MODE, KS64, MBLOCK, 2>(m, d, kap, mas, io, st);
    else launch_valu<double, false, MODE, KS64, MBLOCK, 2>(m, d, kap, mas, io, st);
  } else if (m->mvariant == 1) {
    if (mass) launch_mfma<true, MODE>(m, d, kap, mas, io, st);
    else launch_mfma<false, MODE>(m, d, kap, mas, io, st);
  } else if (m->mvariant == 2) {
    if (mass) launch_valu<float, true, MODE, KS, 1024>(m, d, kap, mas, io, st);
    else launch_valu<float, false, MODE, KS, 1024>(m, d, kap, mas, io, st);
  } else if (m->mvariant == 3) {
    if (mass) launch_valu<float, true, MODE, 4, MBLOCK>(m, d, kap, mas, io, st);
    else launch_valu<float, false, MODE, 4, MBLOCK>(m, d, kap, mas, io, st);
  } else if (m->mvariant == 4) {
    if (mass) launch_valu<float, true, MODE, KS, MBLOCK, 2>(m, d, kap, mas, io, st);
    else launch_valu<float, false, MODE, KS, MBLOCK, 2>(m, d, kap, mas, io, st);
  } else if (m->mvariant == 5) {
    if (mass) launch_valu<float, true, MODE, KS, MBLOCK, 4>(m, d, kap, mas, io, st);
    else launch_valu<float, false, MODE, KS, MBLOCK, 4>(m, d, kap, mas, io, st);
  } else if (m->mvariant == 6) {
    if (mass) launch_valu<float, true, MODE, 4, MBLOCK, 2>(m, d, kap, mas, io, st);
    else launch_valu<float, false, MODE, 4, MBLOCK, 2>(m, d, kap, mas, io, st);
  } else {
    if (mass) launch_valu<float, true, MODE, KS, MBLOCK>(m, d, kap, mas, io, st);
    else launch_valu<float, false, MODE, KS, MBLOCK>(m, d, kap, mas, io, st);
  }
  HIP_CHECK(hipGetLastError());
}

template <class T>
static void restrict_multi(dgh_transfer t, int k, const T* f, T* c, hipStream_t st) {
  const uint32_t nc = t->coarse->N;
  hipLaunchKernelGGL(restrict_multi_kernel<T>, dim3(blocks_for(size_t(nc) * k)), dim3(256), 0, st, nc, k,
                     t->rowptr.p, t->col.p, t->w.p, f, c);
  HIP_CHECK(hipGetLastError());
}

template <class T>
static void prolong_multi(dgh_transfer t, int k, const T* c, const T* f_in, T* f_out, hipStream_t st) {
  const uint32_t nf = t->fine->N;
  hipLaunchKernelGGL(prolong_multi_kernel<T>, dim3(blocks_for(size_t(nf) * k)), dim3(256), 0, st, nf, k,
                     t->celem.p, t->pos.p, t->coarse->nl.p, t->coarse->cs.p, Wtab<T>(), c, f_in, f_out);
  HIP_CHECK(hipGetLastError());
}

template <class T>
static void reduce_multi(uint32_t n, int k, const DotArgs<T>& d, int nq, double* part, double* out,
                         hipStream_t st) {
  hipLaunchKernelGGL(reduce1_multi_kernel<T>, dim3(RED_BLOCKS_M), dim3(256), 0, st, n, k, d, part);
  hipLaunchKernelGGL(reduce2_multi_kernel, dim3(nq * k), dim3(256), 0, st, nq, k, (const double*)part, out, MAXK);
}

// v[:, j] -= mean(v[:, j]) for every column
template <class T>
static void project_mean_multi(dgh_solver s, uint32_t n, int k, T* v, hipStream_t st) {
  DotArgs<T> d{};
  d.a[0] = v;
  reduce_multi<T>(n, k, d, 1, s->mpart_p.p, s->mS.p + S_PROJ * MAXK, st);
  hipLaunchKernelGGL(subtract_mean_multi_kernel<T>, dim3(blocks_for(size_t(n) * k)), dim3(VBLOCK), 0, st, n, k,
                     (const double*)(s->mS.p + S_PROJ * MAXK), v);
}

// Batched FP32 V-cycle on level l (the single right-hand-side vcycle, column by column).
// x1_ready: the first sweep from zero is already in the level's first iterate buffer.
float* mvcycle(dgh_solver s, int l, int k, const float* b, hipStream_t st, bool x1_ready) {
  auto& V = s->m32[l];
  if (l == 0) {
    const int n = s->nc;
    const float* bb = b;
    if (s->prm.null_space) {
      hipLaunchKernelGGL(copy_kernel<float>, dim3(blocks_for(size_t(n) * k)), dim3(VBLOCK), 0, st,
                         uint32_t(size_t(n) * k), b, V.r.p);
      project_mean_multi<float>(s, uint32_t(n), k, V.r.p, st);
      bb = V.r.p;
    }
    hipLaunchKernelGGL(dense_multi_kernel<float>, dim3((n + DM_ROWS - 1) / DM_ROWS, DM_SPLIT), dim3(256), 0, st, n,
                       s->ldc, k, (const float*)s->Ainv32.p, bb, s->mdense.p);
    hipLaunchKernelGGL(dense_multi_sum_kernel<float>, dim3(blocks_for(size_t(n) * k)), dim3(256), 0, st, n, k,
                       (const double*)s->mdense.p, V.xa.p);
    HIP_CHECK(hipGetLastError());
    return V.xa.p;
  }
  dgh_mesh m = s->meshes[l];
  const uint32_t n = m->N;
  const float om = float(s->om_level[l]);
  const float* invd = s->v32[l].invd.p;
  float* xc = V.xa.p;
  float* xn = V.xb.p;
  if (!x1_ready)
    hipLaunchKernelGGL(smooth_first_multi_kernel<float>, dim3(blocks_for(size_t(n) * k)), dim3(VBLOCK), 0, st, n, k,
                       om, b, invd, xc);
  for (int it = 1; it < s->prm.nu_pre; ++it) {
    launch_multi<float, OP_SMOOTH>(m, MultiIO<float>{xc, xn, b, invd, om, k}, st);
    std::swap(xc, xn);
  }
  launch_multi<float, OP_RESID>(m, MultiIO<float>{xc, V.r.p, b, nullptr, 0.f, k}, st);
  restrict_multi<float>(s->transfers[l - 1], k, V.r.p, s->m32[l - 1].b.p, st);
  const float* xcoarse = mvcycle(s, l - 1, k, s->m32[l - 1].b.p, st);
  prolong_multi<float>(s->transfers[l - 1], k, xcoarse, xc, xc, st);
  for (int it = 0; it < s->prm.nu_post; ++it) {
    launch_multi<float, OP_SMOOTH>(m, MultiIO<float>{xc, xn, b, invd, om, k}, st);
    std::swap(xc, xn);
  }
  return xc;
}

// z = B r for k columns.  fused: cg_xr_multi_fused_kernel has already written the FP32
// right-hand side and the first sweep of the finest level.  The result's conversion to FP64
// and the constant-null-space projection are one pass (column means from the FP32 result).
static void mprecondition(dgh_solver s, int k, const double* r, double* z, hipStream_t st, bool fused = false) {
  const int top = s->nlevels - 1;
  const uint32_t n = s->meshes[top]->N;
  const size_t nk = size_t(n) * k;
  auto& V = s->m32[top];
  if (!fused)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, uint32_t(nk), r,
                       V.b.p);
  const float* res = mvcycle(s, top, k, V.b.p, st, fused && top > 0);
  const double* sum = nullptr;
  if (s->prm.null_space) {
    DotArgs<float> d{};
    d.a[0] = res;
    reduce_multi<float>(n, k, d, 1, s->mpart_p.p, s->mS.p + S_PROJ * MAXK, st);
    sum = s->mS.p + S_PROJ * MAXK;
  }
  hipLaunchKernelGGL(convert_sub_mean_multi_kernel, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, n, k, res, sum, z);
}

static void mcg_iteration(dgh_solver s, int k, hipStream_t st) {
  const int top = s->nlevels - 1;
  dgh_mesh m = s->meshes[top];
  const uint32_t n = m->N;
  const size_t nk = size_t(n) * k;
  const bool pnorm = (s->prm.norm_type == DGH_NORM_PRECONDITIONED);
  double* S = s->mS.p;
  hipLaunchKernelGGL(cg_p_multi_kernel, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, n, k, (const double*)S,
                     (const double*)s->mz.p, s->mp.p);
  // w = A p with the tiles' partials of p.w, summed per column in a fixed order
  launch_multi<double, OP_CG>(m, MultiIO<double>{s->mp.p, s->mw.p, nullptr, nullptr, 0.0, k, s->mptile.p}, st);
  const int nt = int(m->mop.plan.tiles.size());
  hipLaunchKernelGGL(sum_tiles_multi_kernel, dim3(k), dim3(256), 0, st, nt, k, (const double*)s->mptile.p,
                     S + S_PW * MAXK);
  auto& V = s->m32[top];
  hipLaunchKernelGGL(cg_xr_multi_fused_kernel, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, n, k, S,
                     (const double*)s->mp.p, (const double*)s->mw.p, s->mx.p, s->mr.p, V.b.p,
                     (const float*)s->v32[top].invd.p, float(s->om_level[top]), V.xa.p);
  mprecondition(s, k, s->mr.p, s->mz.p, st, true);
  DotArgs<double> e{};
  e.a[0] = s->mz.p; e.b[0] = s->mz.p;    // ZZ
  e.a[1] = s->mz.p; e.b[1] = s->mr.p;    // ZR
  e.a[2] = s->mz.p; e.b[2] = s->mw.p;    // ZW (flexible update: the V-cycle runs in FP32)
  e.a[3] = pnorm ? nullptr : s->mr.p; e.b[3] = s->mr.p;   // RR
  reduce_multi<double>(n, k, e, 4, s->mpart.p, S + S_ZZ * MAXK, st);
  hipLaunchKernelGGL(cg_scalar_multi_kernel, dim3(1), dim3(MAXK), 0, st, k, S, int(pnorm), 1, s->prm.rtol);
}

// Sizes the batched buffers of a solver for k columns (FP32 per-level vectors, FP64 outer
// CG vectors, FP32 level diagonals and coarse inverse).
void drop_graphs(dgh_solver s) {
  if (s->mgexec) (void)hipGraphExecDestroy(s->mgexec);
  if (s->mgraph) (void)hipGraphDestroy(s->mgraph);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->mgexec = nullptr;
  s->mgraph = nullptr;
  s->gexec = nullptr;
  s->graph = nullptr;
}

void multi_op_f32(dgh_mesh m, int mode, const float* x, float* y, const float* b, const float* invd, float om,
                  int k, hipStream_t st) {
  const MultiIO<float> io{x, y, b, invd, om, k};
  if (mode == OP_MATVEC) launch_multi<float, OP_MATVEC>(m, io, st);
  else if (mode == OP_RESID) launch_multi<float, OP_RESID>(m, io, st);
  else launch_multi<float, OP_SMOOTH>(m, io, st);
}

void restrict_multi_f32(dgh_transfer t, int k, const float* f, float* c, hipStream_t st) {
  restrict_multi<float>(t, k, f, c, st);
}

void prolong_multi_f32(dgh_transfer t, int k, const float* c, const float* f_in, float* f_out, hipStream_t st) {
  prolong_multi<float>(t, k, c, f_in, f_out, st);
}

void project_mean_multi_f32(dgh_solver s, uint32_t n, int k, float* v, hipStream_t st) {
  project_mean_multi<float>(s, n, k, v, st);
}

void column_dots_f32(uint32_t n, int k, const float* a, const float* b, double* part, double* out, hipStream_t st) {
  DotArgs<float> d{};
  d.a[0] = a;
  d.b[0] = b;
  reduce_multi<float>(n, k, d, 1, part, out, st);
}

void ensure_multi_buffers(dgh_solver s, int k) {
  if (s->coarse_iter)
    fail(DGH_ERR_STATE, "the batched V-cycle needs the dense coarse solve; the coarsest level has %d nodes "
         "(dense limit %d): use more levels", s->nc, std::min(s->prm.coarse_dense_max, COARSE_DENSE_LIMIT));
  if (s->mk == k) return;
  if (s->mgexec) (void)hipGraphExecDestroy(s->mgexec);
  if (s->mgraph) (void)hipGraphDestroy(s->mgraph);
  s->mgexec = nullptr;
  s->mgraph = nullptr;
  const int L = s->nlevels;
  if (s->v32.size() != size_t(L)) {
    s->v32.resize(L);
    for (int l = 0; l < L; ++l) {
      const uint32_t n = s->meshes[l]->N;
      s->v32[l].invd.alloc(n);
      hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n)), dim3(VBLOCK), 0, s->st, n,
                         (const double*)s->v64[l].invd.p, s->v32[l].invd.p);
    }
  }
  if (!s->Ainv32.p) {
    s->Ainv32.alloc(size_t(s->nc) * s->ldc);
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(size_t(s->nc) * s->ldc)), dim3(VBLOCK), 0,
                       s->st, uint32_t(size_t(s->nc) * s->ldc), (const double*)s->Ainv64.p, s->Ainv32.p);
  }
  s->m32.clear();
  s->m32.resize(L);
  for (int l = 0; l < L; ++l) {
    const size_t nk = size_t(s->meshes[l]->N) * k;
    auto& V = s->m32[l];
    V.b.alloc(nk);
    V.xa.alloc(nk);
    V.xb.alloc(nk);
    V.r.alloc(nk);
    if (l > 0) ensure_multi(s->meshes[l]);
  }
  const size_t nk = size_t(s->meshes[L - 1]->N) * k;
  s->mx.alloc(nk);
  s->mr.alloc(nk);
  s->mz.alloc(nk);
  s->mp.alloc(nk);
  s->mw.alloc(nk);
  s->mdense.alloc(size_t(DM_SPLIT) * s->nc * k);
  if (!s->mS.p) {
    s->mS.alloc(size_t(SM_NSCAL) * MAXK);
    s->mpart.alloc(size_t(MAXQ) * RED_BLOCKS_M * MAXK);
    s->mpart_p.alloc(size_t(MAXQ) * RED_BLOCKS_M * MAXK);
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s->hmS), sizeof(double) * SM_NSCAL * MAXK, hipHostMallocDefault));
  }
  HIP_CHECK(hipStreamSynchronize(s->st));
  s->mk = k;
}

}  // namespace dgh

extern "C" {

int dgh_matvec_multi(dgh_mesh m, int prec, int k, const void* x, void* y, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !x || !y) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    check_k(k);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (prec == DGH_F64)
      launch_multi<double, OP_MATVEC>(m, MultiIO<double>{static_cast<const double*>(x), static_cast<double*>(y),
                                                         nullptr, nullptr, 0.0, k}, st);
    else if (prec == DGH_F32)
      launch_multi<float, OP_MATVEC>(m, MultiIO<float>{static_cast<const float*>(x), static_cast<float*>(y),
                                                       nullptr, nullptr, 0.f, k}, st);
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_mesh_set_tile_multi(dgh_mesh m, uint32_t tile_elems, int variant) {
  return guarded([&] {
    require_init();
    if (!m || tile_elems < 32 || variant < 0 || variant >= NVARIANT) fail(DGH_ERR_ARG, "invalid arguments");
    plan_multi(m, tile_elems);
    m->mvariant = variant;
  });
}

int dgh_mesh_autotune_multi(dgh_mesh m, int k, uint32_t* tile_elems, int* variant, float* us) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    check_k(k);
    const size_t nk = size_t(m->N) * k;
    DevBuf<float> xb(nk), yb(nk);
    HIP_CHECK(hipMemset(xb.p, 0, nk * sizeof(float)));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t bt = 0;
    int bv = 0;
    for (uint32_t T : {256u, 512u, 1024u, 2048u}) {
      try {
        plan_multi(m, T);
      } catch (const Error&) {
        continue;
      }
      if (m->mop.plan.tile_elems != T) continue;
      for (int variant = 0; variant < NVARIANT; ++variant) {
        m->mvariant = variant;
        auto run = [&] { launch_multi<float, OP_MATVEC>(m, MultiIO<float>{xb.p, yb.p, nullptr, nullptr, 0.f, k}, nullptr); };
        for (int i = 0; i < 2; ++i) run();
        HIP_CHECK(hipEventRecord(e0, nullptr));
        for (int i = 0; i < 5; ++i) run();
        HIP_CHECK(hipEventRecord(e1, nullptr));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          bt = T;
          bv = variant;
        }
      }
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    if (!bt) fail(DGH_ERR_STATE, "no batched tiling candidate fits");
    plan_multi(m, bt);
    m->mvariant = bv;
    if (tile_elems) *tile_elems = bt;
    if (variant) *variant = bv;
    if (us) *us = best * 200.f;   // ms per 5 launches -> us per launch
  });
}

// KSPSolve_CG per column (the single right-hand-side dgh_solver_solve), k columns at once.
int dgh_solver_solve_multi(dgh_solver s, int k, const double* B, double* X, dgh_solve_stats* stats, void* stream) {
  return guarded([&] {
    require_init();
    if (!s || !B || !X) fail(DGH_ERR_ARG, "NULL argument");
    if (s->prm.coarse_op != DGH_COARSE_JAC2) fail(DGH_ERR_ARG, "the batched solve runs Jacobian2 levels only");
    if (s->dirichlet) fail(DGH_ERR_ARG, "the batched solve has no Dirichlet mode");
    refresh_operators(s);
    check_k(k);
    check_aligned(B, X);
    ensure_multi_buffers(s, k);
    ensure_multi(s->meshes[s->nlevels - 1]);
    const uint64_t gen = mesh_generation(s);
    if (gen != s->graph_gen) {   // re-tiled or new coefficients since the graphs were captured
      drop_graphs(s);
      s->graph_gen = gen;
    }
    // the CG mode's tile partials (sized for the finest level's batched tiling; a re-tiling
    // bumps the mesh generation, which drops the captured graphs above)
    const size_t nptile = s->meshes[s->nlevels - 1]->mop.plan.tiles.size() * size_t(k);
    if (s->mptile.n < nptile) {
      drop_graphs(s);
      s->mptile.alloc(nptile);
    }
    hipStream_t ust = static_cast<hipStream_t>(stream);
    hipStream_t st = s->st;
    const int top = s->nlevels - 1;
    dgh_mesh m = s->meshes[top];
    const uint32_t n = m->N;
    const size_t nk = size_t(n) * k;
    const bool pnorm = (s->prm.norm_type == DGH_NORM_PRECONDITIONED);
    double* S = s->mS.p;
    hipEvent_t e0, e1, eu;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventCreateWithFlags(&eu, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(eu, ust));
    HIP_CHECK(hipStreamWaitEvent(st, eu, 0));
    HIP_CHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, uint32_t(nk), B, s->mr.p);
    if (s->prm.null_space) project_mean_multi<double>(s, n, k, s->mr.p, st);
    hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, s->mx.p, 0.0, uint32_t(nk));
    hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, s->mp.p, 0.0, uint32_t(nk));
    mprecondition(s, k, s->mr.p, s->mz.p, st);
    DotArgs<double> d{};
    d.a[0] = s->mz.p; d.b[0] = s->mz.p;   // ZZ
    d.a[1] = s->mz.p; d.b[1] = s->mr.p;   // ZR
    reduce_multi<double>(n, k, d, 2, s->mpart.p, S + S_ZZ * MAXK, st);
    DotArgs<double> bb{};
    bb.a[0] = s->mr.p; bb.b[0] = s->mr.p;
    reduce_multi<double>(n, k, bb, 1, s->mpart_p.p, S + S_BB * MAXK, st);
    hipLaunchKernelGGL(cg_init_multi_kernel, dim3(1), dim3(MAXK), 0, st, k, S, int(pnorm), s->prm.rtol);
    HIP_CHECK(hipMemcpyAsync(s->hmS, S, sizeof(double) * SM_NSCAL * MAXK, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<double> bnorm(k), dp0(k);
    std::vector<int> its(k, 0), conv(k, 0);
    for (int j = 0; j < k; ++j) {
      bnorm[j] = std::sqrt(s->hmS[S_BB * MAXK + j]);
      dp0[j] = s->hmS[SM_DP0 * MAXK + j];
      conv[j] = s->hmS[SM_ACT * MAXK + j] == 0.0;
    }
    if (s->prm.use_graph && !s->mgexec) {
      HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      mcg_iteration(s, k, st);
      HIP_CHECK(hipStreamEndCapture(st, &s->mgraph));
      HIP_CHECK(hipGraphInstantiate(&s->mgexec, s->mgraph, nullptr, nullptr, 0));
    }
    int it = 0;
    bool bad = false;
    auto all_done = [&] { return std::all_of(conv.begin(), conv.end(), [](int c) { return c != 0; }); };
    while (!all_done() && it < s->prm.max_it) {
      if (s->mgexec)
        HIP_CHECK(hipGraphLaunch(s->mgexec, st));
      else
        mcg_iteration(s, k, st);
      HIP_CHECK(hipMemcpyAsync(s->hmS, S, sizeof(double) * SM_NSCAL * MAXK, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      ++it;
      for (int j = 0; j < k; ++j) {
        if (conv[j]) continue;
        const double dp = s->hmS[S_DP * MAXK + j];
        its[j] = it;
        if (!std::isfinite(dp)) bad = true;
        if (s->hmS[SM_ACT * MAXK + j] == 0.0) conv[j] = 1;
      }
      if (bad) break;
    }
    hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, uint32_t(nk), (const double*)s->mx.p, X);
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (stats) {
      // true residuals ||b_j - A x_j|| / ||b_j|| in FP64
      hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(nk)), dim3(VBLOCK), 0, st, uint32_t(nk), B, s->mw.p);
      if (s->prm.null_space) project_mean_multi<double>(s, n, k, s->mw.p, st);
      launch_multi<double, OP_RESID>(m, MultiIO<double>{X, s->mz.p, s->mw.p, nullptr, 0.0, k}, st);
      DotArgs<double> q{};
      q.a[0] = s->mz.p; q.b[0] = s->mz.p;
      reduce_multi<double>(n, k, q, 1, s->mpart_p.p, S + S_PROJ * MAXK, st);
      HIP_CHECK(hipMemcpyAsync(s->hmS, S, sizeof(double) * SM_NSCAL * MAXK, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      for (int j = 0; j < k; ++j) {
        stats[j].iterations = its[j];
        stats[j].converged = conv[j];
        stats[j].rnorm0 = dp0[j];
        stats[j].rnorm = s->hmS[S_DP * MAXK + j];
        stats[j].rel_true = bnorm[j] > 0 ? std::sqrt(s->hmS[S_PROJ * MAXK + j]) / bnorm[j] : 0.0;
        stats[j].seconds = ms * 1e-3;
      }
    }
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    HIP_CHECK(hipEventDestroy(eu));
    if (bad) fail(DGH_ERR_DIVERGED, "CG breakdown: non-finite residual at iteration %d", it);
  });
}

}  // extern "C"
