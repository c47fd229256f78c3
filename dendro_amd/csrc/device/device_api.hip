// Device side of libdendrohip: mesh upload, operator application, intergrid transfer and
// the GMG-preconditioned CG solver, behind the C ABI of include/dendro_hip.h.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.hpp"

namespace dgh {
State g;

// Tile partials of dgh_matvec_dot_noalloc: a file-scope device array, one copy per device,
// held until the process ends.  In host code `scratch` names a host-side stand-in, not device
// memory: host code uses only the address dgh_init obtains from hipGetSymbolAddress
// (g.noalloc_part[device]).
static __device__ __align__(256) double scratch[DGH_NOALLOC_MAX_TILES];
}  // namespace dgh

namespace dgh {

template <class T>
static size_t op_lds_bytes(const dgh_mesh m) {
  return m->lane_rows == DGH_ROWS_LANE ? tile_lds_bytes<T>(m->op.block, m->op.plan.max_touch, true)
                                       : (sizeof(T) == 8 ? m->op.lds64 : m->op.lds32);
}

// Stencil-row source of a launch: class matrices need uniform coefficients (per-element ones
// fall back to the per-type rows).
static int op_rows_mode(const dgh_mesh m) {
  return (m->lane_rows == DGH_ROWS_CLASS && m->op.rk64.p) ? DGH_ROWS_TYPE : m->lane_rows;
}

// stencil rows of an operator launch: the class matrices for uniform coefficients (unless
// rows are read per lane from LDS), else the child-0 Laplacian and mass rows
template <class T> const T* kctab(const DevTiles& d);
template <> const double* kctab<double>(const DevTiles& d) { return d.kc64.p; }
template <> const float* kctab<float>(const DevTiles& d) { return d.kc32.p; }
template <class T>
static const T* op_rows(dgh_mesh m, bool comb) {
  return comb ? kctab<T>(m->op) : KFtab<T>();
}

template <class T> const T* rk(dgh_mesh m);
template <class T> const T* rm(dgh_mesh m);
template <> const double* rk<double>(dgh_mesh m) { return m->op.rk64.p; }
template <> const double* rm<double>(dgh_mesh m) { return m->op.rm64.p; }
template <> const float* rk<float>(dgh_mesh m) { return m->op.rk32.p; }
template <> const float* rm<float>(dgh_mesh m) { return m->op.rm32.p; }

// The kernel's LDS layout holds op_own_cap(BLOCK) owned nodes: the plan must be the one built
// for that block size.
template <int BLOCK>
static void check_op_plan(dgh_mesh m) {
  if (m->op.block != BLOCK || m->op.plan.own_cap != op_own_cap(BLOCK))
    fail(DGH_ERR_STATE, "operator tiling built for block %d, launched with %d", m->op.block, BLOCK);
}

template <class T, int MODE, int BLOCK>
static void launch_op_b(dgh_mesh m, const OpIO<T>& io, hipStream_t st) {
  check_op_plan<BLOCK>(m);
  const int nt = int(m->op.plan.tiles.size());
  const unsigned grid = unsigned(NXCD * ((nt + NXCD - 1) / NXCD));
  const size_t lds = op_lds_bytes<T>(m);
  const T* kap = rk<T>(m);
  const T* mas = rm<T>(m);
#define DGH_OP_LAUNCH(MASS, UNI, LANE, COMB)                                                                   \
  hipLaunchKernelGGL((tile_op_kernel<T, T, MASS, MODE, BLOCK, UNI, LANE, COMB>), dim3(grid), dim3(BLOCK), lds, st, \
                     m->op.tiles.p, nt, m->op.lid.p, kap, mas, m->lap_fac, m->mass_fac, m->maxD,      \
                     m->hFac, op_rows<T>(m, COMB), m->op.node_tab.p, m->op.texp.p, io)
  const int rows = op_rows_mode(m);
  if (rows == DGH_ROWS_LANE) {
    if (!kap)
      DGH_OP_LAUNCH(true, true, true, false);
    else if (!mas)
      DGH_OP_LAUNCH(false, false, true, false);
    else
      DGH_OP_LAUNCH(true, false, true, false);
  } else if (rows == DGH_ROWS_CLASS) {
    DGH_OP_LAUNCH(true, true, false, true);
  } else {
    if (!kap)
      DGH_OP_LAUNCH(true, true, false, false);
    else if (!mas)
      DGH_OP_LAUNCH(false, false, false, false);
    else
      DGH_OP_LAUNCH(true, false, false, false);
  }
#undef DGH_OP_LAUNCH
  HIP_CHECK(hipGetLastError());
}

template <class T, int MODE>
static void launch_op(dgh_mesh m, const OpIO<T>& io, hipStream_t st) {
  if (m->op.plan.tiles.empty()) return;
  switch (m->block) {
    case 256: launch_op_b<T, MODE, 256>(m, io, st); break;
    case 512: launch_op_b<T, MODE, 512>(m, io, st); break;
    default: launch_op_b<T, MODE, 1024>(m, io, st); break;
  }
}

template <class T>
static void launch_matvec(dgh_mesh m, const T* x, T* y, hipStream_t st) {
  launch_op<T, OP_MATVEC>(m, OpIO<T>{x, y, nullptr, nullptr, T(0)}, st);
}

template <class T>
static void launch_diag(dgh_mesh m, T* d, hipStream_t st) {
  launch_op<T, OP_DIAG>(m, OpIO<T>{nullptr, d, nullptr, nullptr, T(0)}, st);
}

// r = b - A x
template <class T>
static void launch_resid(dgh_mesh m, const T* x, const T* b, T* r, hipStream_t st) {
  launch_op<T, OP_RESID>(m, OpIO<T>{x, r, b, nullptr, T(0)}, st);
}

// x_out = x + omega invd (b - A x)
template <class T>
static void launch_smooth(dgh_mesh m, const T* x, const T* b, const T* invd, T omega, T* x_out, hipStream_t st) {
  launch_op<T, OP_SMOOTH>(m, OpIO<T>{x, x_out, b, invd, omega}, st);
}

// c = R f; with x1 given, also x1 = omega (c * invd) (the first sweep on the coarse level)
// Boundary flags of a transfer's two meshes when both carry them (Dirichlet problem): the
// transfers then drop every entry that touches a boundary node (the suppressed DOFs of
// RestrictMatVec.C:75-80 and ProlongMatVec.C:76-80); nullptrs otherwise.
static inline const uint8_t* tbdy_c(dgh_transfer t) {
  return (t->coarse->bdy.p && t->fine->bdy.p) ? t->coarse->bdy.p : nullptr;
}
static inline const uint8_t* tbdy_f(dgh_transfer t) {
  return (t->coarse->bdy.p && t->fine->bdy.p) ? t->fine->bdy.p : nullptr;
}

template <class T>
static void launch_restrict(dgh_transfer t, const T* f, T* c, hipStream_t st, T omega = T(0),
                            const T* invd = nullptr, T* x1 = nullptr) {
  const uint32_t nc = t->coarse->N;
  if (!nc) return;
  const uint8_t *bc = tbdy_c(t), *bf = tbdy_f(t);
  if (t->ent.p)
    hipLaunchKernelGGL((restrict_kernel<T, true>), dim3(blocks_for(size_t(nc) * RLANES)), dim3(VBLOCK), 0, st, nc,
                       t->rowptr.p, t->ent.p, (const float*)nullptr, f, c, omega, invd, x1, bc, bf);
  else
    hipLaunchKernelGGL((restrict_kernel<T, false>), dim3(blocks_for(size_t(nc) * RLANES)), dim3(VBLOCK), 0, st, nc,
                       t->rowptr.p, t->col.p, t->w.p, f, c, omega, invd, x1, bc, bf);
  HIP_CHECK(hipGetLastError());
}

template <class T>
static void launch_prolong(dgh_transfer t, const T* c, const T* f_in, T* f_out, hipStream_t st) {
  const uint32_t nf = t->fine->N;
  if (!nf) return;
  hipLaunchKernelGGL(prolong_kernel<T>, dim3(blocks_for(nf)), dim3(VBLOCK), 0, st, nf, t->celem.p,
                     t->pos.p, t->coarse->nl.p, t->coarse->cs.p, Wtab<T>(), c, f_in, f_out, tbdy_c(t), tbdy_f(t));
  HIP_CHECK(hipGetLastError());
}

// y = A_D x, the Dirichlet-masked operator (one launch)
template <class T>
static void launch_dirmv(dgh_mesh m, const T* x, T* y, hipStream_t st) {
  launch_op<T, OP_DIRMV>(m, OpIO<T>{x, y, nullptr, nullptr, T(0), nullptr, m->bdy.p}, st);
}

// Jacobian3 on the coarse level of t: y = R (A_f (P x)) with the fine mesh's own coefficients,
// vf and wf fine-level scratch vectors.
template <class T>
static void launch_jac3(dgh_transfer t, const T* x, T* y, T* vf, T* wf, hipStream_t st) {
  launch_prolong<T>(t, x, nullptr, vf, st);
  launch_matvec<T>(t->fine, vf, wf, st);
  launch_restrict<T>(t, wf, y, st);
}

// Per-record element factors of a tiling (they follow its record order), and the per-tile
// bounds of the fixed-point sums: texp[t].x = e with
//   |fac1| max_k sum_j |L_t[k][j]| + |fac2| max_k sum_j |M_t[k][j]| < 2^e
// over the tile's records, factors formed as the kernel forms them (elem_factors), and
// texp[t].y = v with every owned node of the tile receiving at most 2^v terms.
void upload_coefs(dgh_mesh m, DevTiles& d) {
  ++m->gen;
  d.rk64.release();
  d.rm64.release();
  d.rk32.release();
  d.rm32.release();
  const size_t R = d.plan.rec_elem.size();
  if (R == 0) return;
  const bool per_elem = !m->h_kappa.empty();
  const bool mass = per_elem ? !m->h_mass.empty() : true;
  std::vector<int2> te(d.plan.tiles.size());
  std::vector<uint32_t> terms;
  for (size_t t = 0; t < d.plan.tiles.size(); ++t) {
    const TileRec& tr = d.plan.tiles[t];
    terms.assign(tr.n_own, 0u);
    for (size_t q = size_t(tr.elem_off) * 8; q < size_t(tr.elem_off + tr.n_elem) * 8; ++q)
      if (d.plan.lid[q] < tr.n_own) terms[d.plan.lid[q]]++;
    uint32_t vmax = 1;
    for (uint32_t c : terms) vmax = std::max(vmax, c);
    int v = 0;
    while ((1u << v) < vmax) ++v;
    double mx = 0.0;
    for (uint32_t r = tr.elem_off; r < tr.elem_off + tr.n_elem; ++r) {
      const uint32_t e = d.plan.rec_elem[r];
      const double h = m->hFac * double(1u << (m->maxD - m->h_lev[e]));
      const double k1 = per_elem ? m->h_kappa[e] : m->lap_fac;
      const double k2 = per_elem ? (mass ? m->h_mass[e] : 0.0) : m->mass_fac;
      const int tp = d.plan.meta[r] & 31;
      mx = std::max(mx, std::fabs(k1 * h / 2.0) * g.lrow[tp] + (mass ? std::fabs(k2 * h * h * h / 8.0) * g.mrow[tp] : 0.0));
    }
    if (!std::isfinite(mx)) fail(DGH_ERR_ARG, "non-finite element coefficients");
    int e = -1022;
    if (mx > 0.0) std::frexp(mx, &e);   // mx < 2^e
    te[t] = make_int2(std::max(e, -1022), v);
  }
  d.texp.upload(te.data(), te.size());
  d.kc64.release();
  d.kc32.release();
  if (!per_elem) {
    // element matrices per record class, K = fac1 L + fac2 M with the factors elem_factors
    // forms (each entry rounded once)
    std::vector<double> kc(size_t(NCLASS) * 64, 0.0);
    std::vector<uint8_t> seen(NCLASS, 0);
    for (size_t r = 0; r < R; ++r) {
      const uint32_t c = d.plan.meta[r];
      if (c >= uint32_t(NCLASS)) fail(DGH_ERR_MESH, "record class %u out of range", c);
      if (seen[c]) continue;
      seen[c] = 1;
      const double h = m->hFac * double(1u << (m->maxD - (c >> 5)));
      const double f1 = m->lap_fac * h / 2.0, f2 = m->mass_fac * h * h * h / 8.0;
      const double* KF = g.hKF.data() + size_t(c & 31u) * KFSTRIDE;
      for (int q = 0; q < 64; ++q) kc[size_t(c) * 64 + q] = f1 * KF[q] + f2 * KF[64 + q];
    }
    d.kc64.upload(kc.data(), kc.size());
    std::vector<float> kcf(kc.begin(), kc.end());
    d.kc32.upload(kcf.data(), kcf.size());
    return;
  }
  std::vector<double> k(R);
  std::vector<float> kf(R);
  // per-record element factors, formed as elem_factors forms them
  auto h_of = [&](size_t r) { return m->hFac * double(1u << (m->maxD - m->h_lev[d.plan.rec_elem[r]])); };
  for (size_t r = 0; r < R; ++r) {
    const double h = h_of(r);
    kf[r] = float(k[r] = m->h_kappa[d.plan.rec_elem[r]] * h / 2.0);
  }
  d.rk64.upload(k.data(), R);
  d.rk32.upload(kf.data(), R);
  if (!m->h_mass.empty()) {
    for (size_t r = 0; r < R; ++r) {
      const double h = h_of(r);
      kf[r] = float(k[r] = m->h_mass[d.plan.rec_elem[r]] * h * h * h / 8.0);
    }
    d.rm64.upload(k.data(), R);
    d.rm32.upload(kf.data(), R);
  }
}

// The device form of a plan's records (load_rec): each record's 8 vertex ids with its class
// (child-0 type | level << 5) in the top 2 bits of ids 0..4, so no class array is read.
static std::vector<uint16_t> pack_records(const TilePlan& plan) {
  if (plan.max_touch > REC_MAX_TOUCH)
    fail(DGH_ERR_ARG, "a tile touches %u nodes (at most %u): use a smaller tile size", plan.max_touch, REC_MAX_TOUCH);
  std::vector<uint16_t> p(plan.lid);
  for (size_t r = 0; r < plan.meta.size(); ++r) {
    const uint32_t c = plan.meta[r];
    if (c >= 1024u) fail(DGH_ERR_MESH, "record class %u out of range", c);
    for (int q = 0; q < 5; ++q) p[r * 8 + q] = uint16_t(p[r * 8 + q] | (((c >> (2 * q)) & 3u) << 14));
  }
  return p;
}

// (Re)builds a tiling with the largest tile size <= T whose LDS need (lds(d)) fits; with
// budget > 0 tiles also end before their owned plus touched nodes exceed the budget, and with
// block > 0 (single right-hand-side operator) the owned nodes get the fixed capacity
// op_own_cap(block) of the kernel's LDS layout.
void build_devtiles(dgh_mesh m, DevTiles& d, uint32_t T, size_t (*lds)(const DevTiles&), uint32_t budget,
                    int block) {
  ++m->gen;
  const uint32_t own_cap = block > 0 ? op_own_cap(block) : 0u;
  for (;;) {
    build_tiles(m->E, m->N, m->h_nl.data(), m->h_child.data(), m->h_hn.data(), m->h_lev.data(), T, d.plan, budget,
                own_cap);
    d.ys_len = (d.plan.max_own + 1) & ~1u;
    d.block = block;
    d.lds64 = block > 0 ? uint32_t(tile_lds_bytes<double>(block, d.plan.max_touch)) : 0u;
    d.lds32 = block > 0 ? uint32_t(tile_lds_bytes<float>(block, d.plan.max_touch)) : 0u;
    if (lds(d) <= size_t(g.max_lds) || T <= 32) break;
    T /= 2;
  }
  if (lds(d) > size_t(g.max_lds))
    fail(DGH_ERR_ARG, "operator tiles need %zu bytes of LDS (limit %d)", lds(d), g.max_lds);
  d.tiles.upload(d.plan.tiles.data(), d.plan.tiles.size());
  const std::vector<uint16_t> packed = pack_records(d.plan);
  d.lid.upload(packed.data(), packed.size());
  d.node_tab.upload(d.plan.node_tab.data(), d.plan.node_tab.size());
  upload_coefs(m, d);
}

static size_t op_lds_f64(const DevTiles& d) { return d.lds64; }

// Node budget of a single right-hand-side tile for a block size: the LDS share that lets the
// wave limit (32 wavefronts per CU) rather than LDS bound the workgroups per CU, less a
// 512-byte allowance for static LDS and allocation granularity.
static uint32_t op_budget(int block) {
  const int per_cu = std::max(1, 2048 / block);
  return uint32_t((OP_LDS_PER_CU / per_cu - 512) / 8);
}

// (Re)builds the single right-hand-side tiling with tiles of at most T elements within the
// block size's node budget.
static void plan_mesh(dgh_mesh m, uint32_t T) {
  build_devtiles(m, m->op, T, op_lds_f64, op_budget(m->block), m->block);
  if (m->rows_auto)   // class matrices once the tiles fill the device several times over
    m->lane_rows = m->op.plan.tiles.size() >= CLASS_ROWS_MIN_TILES ? DGH_ROWS_CLASS : DGH_ROWS_TYPE;
  if (m->lane_rows == DGH_ROWS_LANE && tile_lds_bytes<double>(m->block, m->op.plan.max_touch, true) > size_t(g.max_lds))
    m->lane_rows = DGH_ROWS_TYPE;   // no LDS room left for the stencil table
}

}  // namespace dgh

// ---------------------------------------------------------------- solver
namespace dgh {

template <class T>
static void reduce(uint32_t n, const DotArgs<T>& d, int nq, double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL(reduce1_kernel<T>, dim3(RED_BLOCKS), dim3(VBLOCK), 0, st, n, d, part);
  hipLaunchKernelGGL(reduce2_kernel, dim3(1), dim3(RED2_THREADS), 0, st, nq, (const double*)part, out);
}

// v -= mean(v)
template <class T>
static void project_mean(dgh_solver s, uint32_t n, T* v, hipStream_t st) {
  DotArgs<T> d{};
  d.a[0] = v;
  reduce<T>(n, d, 1, s->part_p.p, s->S.p + S_PROJ, st);
  hipLaunchKernelGGL(subtract_mean_kernel<T>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n,
                     (const double*)(s->S.p + S_PROJ), v);
}

template <class T> std::vector<LevelVecs<T>>& vecs(dgh_solver s);
template <> std::vector<LevelVecs<double>>& vecs<double>(dgh_solver s) { return s->v64; }
template <> std::vector<LevelVecs<float>>& vecs<float>(dgh_solver s) { return s->v32; }
template <class T> const T* ainvp(dgh_solver s);
template <> const double* ainvp<double>(dgh_solver s) { return s->Ap64.p; }
template <> const float* ainvp<float>(dgh_solver s) { return s->Ap32.p; }

// The coarsest level's operator in the iterative coarse solve: the operator the dense path
// assembles (Jacobian3: R A_1 P with level 1's coefficients; Dirichlet: the masked operator).
static void coarse_op_apply(dgh_solver s, const double* x, double* y, hipStream_t st) {
  dgh_mesh m = s->meshes[0];
  if (s->prm.coarse_op == DGH_COARSE_JAC3 && s->nlevels > 1)
    launch_jac3<double>(s->transfers[0], x, y, s->v64[1].r.p, s->v64[1].t.p, st);
  else if (s->dirichlet)
    launch_dirmv<double>(m, x, y, st);
  else
    launch_matvec<double>(m, x, y, st);
}

// x = A^-1 b on a coarsest level above the dense limit: Jacobi-preconditioned CG in FP64 from
// a zero guess to ||b - A x|| <= coarse_rtol ||b||, in chunks of CCG_CHUNK iterations with the
// device-side stopping test read back after each chunk (so this path does not run under
// graph capture).  With a constant null space b is projected first and x afterwards: the
// mean-free solution, as the dense path's (A + a 11^T)^-1 gives for a mean-free b.
constexpr int CCG_CHUNK = 16;
template <class T>
static void coarse_cg(dgh_solver s, const T* b, T* x, hipStream_t st) {
  const uint32_t n = uint32_t(s->nc);
  const dim3 gb(blocks_for(n)), tb(VBLOCK);
  double* S = s->cS.p;
  const double* invd = s->v64[0].invd.p;
  hipLaunchKernelGGL(ccg_load_kernel<T>, gb, tb, 0, st, n, b, s->cr.p);
  if (s->prm.null_space) project_mean<double>(s, n, s->cr.p, st);
  hipLaunchKernelGGL(ccg_start_kernel, gb, tb, 0, st, n, invd, (const double*)s->cr.p, s->cx.p, s->cz.p, s->cp.p);
  DotArgs<double> rz{};
  rz.a[0] = s->cr.p; rz.b[0] = s->cz.p;   // CC_RZN
  rz.a[1] = s->cr.p; rz.b[1] = s->cr.p;   // CC_RR
  DotArgs<double> pw{};
  pw.a[0] = s->cp.p; pw.b[0] = s->cw.p;
  reduce<double>(n, rz, 2, s->part.p, S + CC_RZN, st);
  hipLaunchKernelGGL(ccg_init_kernel, dim3(1), dim3(1), 0, st, S, s->prm.coarse_rtol * s->prm.coarse_rtol);
  int its = 0;
  for (;;) {
    for (int q = 0; q < CCG_CHUNK; ++q) {
      coarse_op_apply(s, s->cp.p, s->cw.p, st);
      reduce<double>(n, pw, 1, s->part.p, S + CC_PW, st);
      hipLaunchKernelGGL(ccg_alpha_kernel, dim3(1), dim3(1), 0, st, S);
      hipLaunchKernelGGL(ccg_update_kernel, gb, tb, 0, st, n, (const double*)S, (const double*)s->cp.p,
                         (const double*)s->cw.p, invd, s->cx.p, s->cr.p, s->cz.p);
      reduce<double>(n, rz, 2, s->part.p, S + CC_RZN, st);
      hipLaunchKernelGGL(ccg_beta_kernel, dim3(1), dim3(1), 0, st, S);
      hipLaunchKernelGGL(ccg_p_kernel, gb, tb, 0, st, n, (const double*)S, (const double*)s->cz.p, s->cp.p);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(s->hcS, S, sizeof(double) * CC_N, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    its = int(s->hcS[CC_ITS]);
    if (s->hcS[CC_DONE] != 0.0 || its >= s->prm.coarse_max_it) break;
  }
  s->coarse_its = its;
  if (s->prm.null_space) project_mean<double>(s, n, s->cx.p, st);
  hipLaunchKernelGGL((convert_kernel<T, double>), gb, tb, 0, st, n, (const double*)s->cx.p, x);
  HIP_CHECK(hipGetLastError());
}

// x = Ainv b on the coarsest level (packed symmetric blocks, two passes), or the iterative
// coarse solve above the dense limit
template <class T>
static void coarse_apply(dgh_solver s, const T* b, T* x, hipStream_t st) {
  if (s->coarse_iter) {
    coarse_cg<T>(s, b, x, st);
    return;
  }
  const int nblk = s->nbc * (s->nbc + 1) / 2;
  hipLaunchKernelGGL(sym_block_kernel<T>, dim3(nblk), dim3(128), 0, st, s->nc, s->npc, ainvp<T>(s), b, s->Qc.p);
  hipLaunchKernelGGL(sym_sum_kernel<T>, dim3(s->nbc), dim3(256), 0, st, s->nc, s->nbc, s->npc,
                     (const double*)s->Qc.p, b, (const double*)(s->prm.null_space ? s->c1.p : nullptr), x);
}

// PCMG multiplicative V-cycle on level l for right-hand side b (PCMGMCycle_Private with the
// options of bin/options:182-206): pre-smoothing from a zero guess, residual, restriction,
// recursion (coarsest: direct solve), interpolate-add, post-smoothing.  Returns the buffer
// holding the result (one of the level's two iterate buffers).  The first sweep from zero,
// x = omega (b * invd), needs no operator: with x1_ready the kernel that produced b has
// already written it into the level's first iterate buffer (the restriction does so for
// every level below the top).
template <class T>
static T* vcycle(dgh_solver s, int l, const T* b, hipStream_t st, bool x1_ready = false) {
  auto& VV = vecs<T>(s);
  auto& V = VV[l];
  if (l == 0) {
    coarse_apply<T>(s, b, V.xa.p, st);   // projects b first with a constant null space
    return V.xa.p;
  }
  dgh_mesh m = s->meshes[l];
  const uint32_t n = m->N;
  const T om = T(s->om_level[l]);
  T* xc = V.xa.p;
  T* xn = V.xb.p;
  // Jacobian3 below the finest level: the level's operator is R A_{l+1} P, applied into V.r
  // with the next finer level's r and t as scratch (its r is free once restricted); the
  // Jacobi diagonal stays the level's own
  const bool j3 = s->prm.coarse_op == DGH_COARSE_JAC3 && l < s->nlevels - 1;
  // Dirichlet problem: the masked operator (one launch into V.r), then the sweep or residual
  auto apply = [&](const T* xi) {
    if (j3)
      launch_jac3<T>(s->transfers[l], xi, V.r.p, VV[l + 1].r.p, VV[l + 1].t.p, st);
    else
      launch_dirmv<T>(m, xi, V.r.p, st);
  };
  const bool split = j3 || s->dirichlet;
  auto sweep = [&](const T* xi, T* xo) {
    if (!split) {
      launch_smooth<T>(m, xi, b, V.invd.p, om, xo, st);
      return;
    }
    apply(xi);
    hipLaunchKernelGGL(smooth_vec_kernel<T>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, xi, b, (const T*)V.r.p,
                       (const T*)V.invd.p, om, xo);
  };
  if (!x1_ready)
    hipLaunchKernelGGL(smooth_first_kernel<T>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, om, b,
                       (const T*)V.invd.p, xc);
  for (int it = 1; it < s->prm.nu_pre; ++it) {
    sweep(xc, xn);
    std::swap(xc, xn);
  }
  if (split) {
    apply(xc);
    hipLaunchKernelGGL(resid_vec_kernel<T>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, b, (const T*)V.r.p, V.r.p);
  } else {
    launch_resid<T>(m, xc, b, V.r.p, st);
  }
  auto& C = VV[l - 1];
  if (l - 1 > 0 && s->prm.nu_pre > 0)
    launch_restrict<T>(s->transfers[l - 1], V.r.p, C.b.p, st, T(s->om_level[l - 1]), C.invd.p, C.xa.p);
  else
    launch_restrict<T>(s->transfers[l - 1], V.r.p, C.b.p, st);
  T* xcoarse = vcycle<T>(s, l - 1, C.b.p, st, l - 1 > 0 && s->prm.nu_pre > 0);
  launch_prolong<T>(s->transfers[l - 1], xcoarse, xc, xc, st);
  for (int it = 0; it < s->prm.nu_post; ++it) {
    sweep(xc, xn);
    std::swap(xc, xn);
  }
  return xc;
}

// z = B r at the finest level, FP64 in and out.
static void precondition(dgh_solver s, const double* r, double* z, hipStream_t st) {
  const int top = s->nlevels - 1;
  const uint32_t n = s->meshes[top]->N;
  if (s->prm.vcycle_prec == DGH_F32) {
    auto& V = s->v32[top];
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, r, V.b.p);
    const float* res = vcycle<float>(s, top, V.b.p, st);
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, res, z);
  } else {
    const double* res = vcycle<double>(s, top, r, st);
    hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, res, z);
  }
  if (s->prm.null_space) project_mean<double>(s, n, z, st);
}

// z = B r on the finest level, unprojected, in the V-cycle's precision (the buffer is the
// same on every call).  Mixed: the FP32 copy of r must already sit in the level's b vector.
static const void* precondition_raw(dgh_solver s, const double* r, hipStream_t st, bool x1_ready = false) {
  const int top = s->nlevels - 1;
  if (s->prm.vcycle_prec == DGH_F32) return vcycle<float>(s, top, s->v32[top].b.p, st, x1_ready);
  return vcycle<double>(s, top, r, st, x1_ready);
}

template <int BLOCK>
static void launch_cg_b(dgh_mesh m, const double* p, double* w, double* part, hipStream_t st) {
  check_op_plan<BLOCK>(m);
  const int nt = int(m->op.plan.tiles.size());
  const unsigned grid = unsigned(NXCD * ((nt + NXCD - 1) / NXCD));
  const size_t lds = op_lds_bytes<double>(m);
  const OpIO<double> io{p, w, nullptr, nullptr, 0.0, part};
  const double* kap = m->op.rk64.p;
  const double* mas = m->op.rm64.p;
#define DGH_CG_LAUNCH(MASS, UNI, LANE, COMB)                                                                    \
  hipLaunchKernelGGL((tile_op_kernel<double, double, MASS, OP_CG, BLOCK, UNI, LANE, COMB>), dim3(grid), dim3(BLOCK), \
                     lds, st, m->op.tiles.p, nt, m->op.lid.p, kap, mas, m->lap_fac, m->mass_fac,      \
                     m->maxD, m->hFac, op_rows<double>(m, COMB), m->op.node_tab.p, m->op.texp.p, io)
  const int rows = op_rows_mode(m);
  if (rows == DGH_ROWS_LANE) {
    if (!kap)
      DGH_CG_LAUNCH(true, true, true, false);
    else if (!mas)
      DGH_CG_LAUNCH(false, false, true, false);
    else
      DGH_CG_LAUNCH(true, false, true, false);
  } else if (rows == DGH_ROWS_CLASS) {
    DGH_CG_LAUNCH(true, true, false, true);
  } else {
    if (!kap)
      DGH_CG_LAUNCH(true, true, false, false);
    else if (!mas)
      DGH_CG_LAUNCH(false, false, false, false);
    else
      DGH_CG_LAUNCH(true, false, false, false);
  }
#undef DGH_CG_LAUNCH
  HIP_CHECK(hipGetLastError());
}

// w = A p and the tile partials of p.w -> part
static void launch_cg(dgh_mesh m, const double* p, double* w, double* part, hipStream_t st) {
  switch (m->block) {
    case 256: launch_cg_b<256>(m, p, w, part, st); break;
    case 512: launch_cg_b<512>(m, p, w, part, st); break;
    default: launch_cg_b<1024>(m, p, w, part, st); break;
  }
}

// One CG iteration (everything between two stopping tests), on device scalars only:
// direction update (projected z), the operator with the p.w tile partials, alpha, x and r
// updates (and the FP32 copy of r), the V-cycle, one pass for all dot products (z kept
// unprojected, see cg_dots_kernel), the scalars.
static void cg_iteration(dgh_solver s, hipStream_t st) {
  const int top = s->nlevels - 1;
  dgh_mesh m = s->meshes[top];
  const uint32_t n = m->N;
  const bool mixed = (s->prm.vcycle_prec == DGH_F32);
  const bool pnorm = (s->prm.norm_type == DGH_NORM_PRECONDITIONED);
  double* S = s->S.p;
  if (mixed)
    hipLaunchKernelGGL(cg_p_fused_kernel<float>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, (const double*)S,
                       static_cast<const float*>(s->zres), s->p.p);
  else
    hipLaunchKernelGGL(cg_p_fused_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, (const double*)S,
                       static_cast<const double*>(s->zres), s->p.p);
  // alpha from the partials of p.w: the operator's tile partials, or (Dirichlet problem) the
  // block partials of a separate dot pass
  int nt = int(m->op.plan.tiles.size());
  if (s->dirichlet) {
    launch_dirmv<double>(m, s->p.p, s->w.p, st);
    DotArgs<double> q{};
    q.a[0] = s->p.p;
    q.b[0] = s->w.p;
    hipLaunchKernelGGL(reduce1_kernel<double>, dim3(RED_BLOCKS), dim3(VBLOCK), 0, st, n, q, s->ptile.p);
    nt = RED_BLOCKS;
  } else {
    launch_cg(m, s->p.p, s->w.p, s->ptile.p, st);
  }
  const bool x1 = s->nlevels > 1 && s->prm.nu_pre > 0;   // first sweep of the V-cycle fused
  hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(RED2_THREADS), 0, st, nt, (const double*)s->ptile.p, S);
  if (mixed)
    hipLaunchKernelGGL(cg_xr_fused_kernel<float>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, (const double*)S,
                       (const double*)s->p.p, (const double*)s->w.p, s->x.p, s->r.p, s->v32[top].b.p,
                       float(s->om_level[top]), (const float*)s->v32[top].invd.p,
                       x1 ? s->v32[top].xa.p : s->v32[top].xb.p);
  else
    hipLaunchKernelGGL(cg_xr_fused_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, (const double*)S,
                       (const double*)s->p.p, (const double*)s->w.p, s->x.p, s->r.p, (float*)nullptr,
                       s->om_level[top], (const double*)s->v64[top].invd.p, x1 ? s->v64[top].xa.p : s->v64[top].xb.p);
  const void* z = precondition_raw(s, s->r.p, st, x1);
  if (z != s->zres) fail(DGH_ERR_STATE, "V-cycle result buffer changed");
  if (mixed)
    hipLaunchKernelGGL(cg_dots_kernel<float>, dim3(DOT_BLOCKS), dim3(VBLOCK), 0, st, n, static_cast<const float*>(z),
                       (const double*)s->r.p, (const double*)s->w.p, s->part.p);
  else
    hipLaunchKernelGGL(cg_dots_kernel<double>, dim3(DOT_BLOCKS), dim3(VBLOCK), 0, st, n, static_cast<const double*>(z),
                       (const double*)s->r.p, (const double*)s->w.p, s->part.p);
  hipLaunchKernelGGL(cg_scalar_fused_kernel, dim3(1), dim3(RED2_THREADS), 0, st, (const double*)s->part.p, n,
                     int(s->prm.null_space), S, int(pnorm), int(mixed));
}

// Jacobi diagonal of level l (the Dirichlet one, 1 on boundary nodes, for the Dirichlet problem)
static void solver_diag(dgh_solver s, int l, double* d) {
  dgh_mesh m = s->meshes[l];
  launch_op<double, OP_DIAG>(m, OpIO<double>{nullptr, d, nullptr, nullptr, 0.0, nullptr,
                                             s->dirichlet ? m->bdy.p : nullptr}, s->st);
}

// Jacobian3 coarsest matrix (BUILD_FULL_JAC_TYPE3_BLOCK, examples/src/backend/omgJac.C:2341-2448)
// as the one-level Galerkin product P^T A_f P, to which its Type2/Type1 element blocks sum:
// every element e of the fine mesh adds (S_e P)^T K_e (S_e P), where row k of S_e P holds the
// coarse shape-function weights at the node of vertex k (the prolongation's own weights).
static void assemble_jac3_coarse(dgh_transfer t, const double* L, const double* M, double* A, int ld) {
  dgh_mesh cm = t->coarse, fm = t->fine;
  std::vector<double> W(144 * 27 * 8);
  generate_transfer_weights(W.data());
  const bool per = !fm->h_kappa.empty();
  for (uint32_t e = 0; e < fm->E; ++e) {
    const double h = fm->hFac * double(1u << (fm->maxD - fm->h_lev[e]));
    const double k1 = per ? fm->h_kappa[e] : fm->lap_fac;
    const double k2 = per ? (fm->h_mass.empty() ? 0.0 : fm->h_mass[e]) : fm->mass_fac;
    const double fac1 = k1 * h / 2.0, fac2 = k2 * h * h * h / 8.0;
    const double* Ls = L + size_t(fm->h_cs[e]) * 64;
    const double* Ms = M + size_t(fm->h_cs[e]) * 64;
    uint32_t cn[8][8];
    double cw[8][8];
    int cnt[8];
    for (int k = 0; k < 8; ++k) {
      const uint32_t nf = fm->h_nl[size_t(e) * 8 + k];
      const uint32_t ce = t->h_celem[nf];
      const double* w = W.data() + (size_t(cm->h_cs[ce]) * 27 + t->h_pos[nf]) * 8;
      cnt[k] = 0;
      for (int q = 0; q < 8; ++q)
        if (w[q] != 0.0) {
          cn[k][cnt[k]] = cm->h_nl[size_t(ce) * 8 + q];
          cw[k][cnt[k]++] = w[q];
        }
    }
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 8; ++j) {
        const double kv = fac1 * Ls[k * 8 + j] + fac2 * Ms[k * 8 + j];
        for (int a = 0; a < cnt[k]; ++a) {
          double* row = A + size_t(cn[k][a]) * ld;
          const double ka = kv * cw[k][a];
          for (int b = 0; b < cnt[j]; ++b) row[cn[j][b]] += ka * cw[j][b];
        }
      }
  }
}

// Assembles the coarsest operator from its element blocks and inverts it on the device
// (Cholesky, rocSOLVER).  With a constant null space, A + a 11^T (a = trace / n^2).
static void setup_coarse(dgh_solver s) {
  dgh_mesh m = s->meshes[0];
  const int n = int(m->N);
  s->nc = n;
  if (n > std::min(s->prm.coarse_dense_max, COARSE_DENSE_LIMIT)) {   // iterative coarse solve
    s->coarse_iter = true;
    if (!s->cx.p) {
      s->cx.alloc(n);
      s->cr.alloc(n);
      s->cz.alloc(n);
      s->cp.alloc(n);
      s->cw.alloc(n);
      s->cS.alloc(CC_N);
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s->hcS), sizeof(double) * CC_N, hipHostMallocDefault));
    }
    return;
  }
  std::vector<double> L(9216), M(9216);
  generate_stencils(L.data(), M.data(), nullptr, nullptr);
  const int ld = (n + DENSE_PAD - 1) / DENSE_PAD * DENSE_PAD;
  s->ldc = ld;
  std::vector<double> A(size_t(n) * ld, 0.0);
  if (s->prm.coarse_op == DGH_COARSE_JAC3 && s->nlevels > 1) {
    assemble_jac3_coarse(s->transfers[0], L.data(), M.data(), A.data(), ld);
  } else {
  const double hFac = m->hFac;
  for (uint32_t e = 0; e < m->E; ++e) {
    const double h = hFac * double(1u << (m->maxD - m->h_lev[e]));
    const double k1 = m->h_kappa.empty() ? m->lap_fac : m->h_kappa[e];
    const double k2 = m->h_kappa.empty() ? m->mass_fac : (m->h_mass.empty() ? 0.0 : m->h_mass[e]);
    const double fac1 = k1 * h / 2.0, fac2 = k2 * h * h * h / 8.0;
    const double* Ls = L.data() + size_t(m->h_cs[e]) * 64;
    const double* Ms = M.data() + size_t(m->h_cs[e]) * 64;
    const uint32_t* idx = m->h_nl.data() + size_t(e) * 8;
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 8; ++j) A[size_t(idx[k]) * ld + idx[j]] += fac1 * Ls[k * 8 + j] + fac2 * Ms[k * 8 + j];
  }
  }
  if (s->dirichlet) {   // ComputeDirichletJacobian (omgJac.C:382-506): identity rows and columns
    const uint8_t* bd = m->h_bdy.data();
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (bd[i] || bd[j]) A[size_t(i) * ld + j] = (i == j) ? 1.0 : 0.0;
  }
  if (s->prm.null_space) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += A[size_t(i) * ld + i];
    const double a = tr / (double(n) * n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) A[size_t(i) * ld + j] += a;
  }
  s->Ainv64.upload(A.data(), A.size());
  if (!g.blas) {
    if (rocblas_create_handle(&g.blas) != rocblas_status_success) fail(DGH_ERR_HIP, "rocblas_create_handle failed");
  }
  DevBuf<rocblas_int> info(1);
  if (rocsolver_dpotrf(g.blas, rocblas_fill_lower, n, s->Ainv64.p, ld, info.p) != rocblas_status_success)
    fail(DGH_ERR_HIP, "rocsolver_dpotrf failed");
  rocblas_int hinfo = 0;
  HIP_CHECK(hipMemcpy(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost));
  if (hinfo != 0) fail(DGH_ERR_DIVERGED, "coarsest operator is not positive definite (pivot %d)", int(hinfo));
  if (rocsolver_dpotri(g.blas, rocblas_fill_lower, n, s->Ainv64.p, ld, info.p) != rocblas_status_success)
    fail(DGH_ERR_HIP, "rocsolver_dpotri failed");
  HIP_CHECK(hipMemcpy(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost));
  if (hinfo != 0) fail(DGH_ERR_DIVERGED, "coarsest inverse failed (%d)", int(hinfo));
  hipLaunchKernelGGL(symmetrize_kernel<double>, dim3((n + 255) / 256, n), dim3(256), 0, 0, n, ld, s->Ainv64.p);
  HIP_CHECK(hipGetLastError());
  s->nbc = (n + SB - 1) / SB;
  s->npc = s->nbc * SB;
  const int nblk = s->nbc * (s->nbc + 1) / 2;
  s->Ap64.alloc(size_t(nblk) * SB * SB);
  s->Qc.alloc(size_t(s->nbc) * s->npc);
  hipLaunchKernelGGL(sym_pack_kernel<double>, dim3(nblk), dim3(256), 0, 0, n, ld, (const double*)s->Ainv64.p, s->Ap64.p);
  HIP_CHECK(hipGetLastError());
  if (s->prm.null_space) {   // Ainv 1, for the projection folded into the coarse solve
    s->c1.alloc(n);
    hipLaunchKernelGGL(row_sum_kernel<double>, dim3((n + 3) / 4), dim3(256), 0, 0, n, ld, (const double*)s->Ainv64.p,
                       s->c1.p);
    HIP_CHECK(hipGetLastError());
  }
  if (s->prm.vcycle_prec == DGH_F32) {
    s->Ainv32.alloc(size_t(n) * ld);
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(size_t(n) * ld)), dim3(VBLOCK), 0, 0,
                       uint32_t(size_t(n) * ld), (const double*)s->Ainv64.p, s->Ainv32.p);
    s->Ap32.alloc(size_t(nblk) * SB * SB);
    hipLaunchKernelGGL(sym_pack_kernel<float>, dim3(nblk), dim3(256), 0, 0, n, ld, (const float*)s->Ainv32.p, s->Ap32.p);
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipDeviceSynchronize());
}

// After dgh_mesh_set_coefficients on a solver's mesh: the level's Jacobi diagonal (and, for
// the coarsest level or the level a Jacobian3 coarsest matrix is built from, the coarsest
// inverse) are rebuilt, and learned weights of that level give way to plain Jacobi.  Boundary
// flags cannot change under a solver: the Dirichlet mode and the masked coarsest matrix were
// fixed at dgh_solver_create.
void refresh_operators(dgh_solver s) {
  bool coarse = false;
  for (int l = 0; l < s->nlevels; ++l) {
    dgh_mesh m = s->meshes[l];
    if (m->bdy_gen != s->bdy_gen[l])
      fail(DGH_ERR_STATE, "boundary flags of level %d changed after dgh_solver_create", l);
    if (m->coef_gen == s->coef_gen[l]) continue;
    s->coef_gen[l] = m->coef_gen;
    const uint32_t n = m->N;
    auto& V = s->v64[l];
    solver_diag(s, l, V.r.p);
    hipLaunchKernelGGL(reciprocal_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, s->st, n,
                       (const double*)V.r.p, V.invd.p);
    if (!s->v32.empty())
      hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n)), dim3(VBLOCK), 0, s->st, n,
                         (const double*)V.invd.p, s->v32[l].invd.p);
    HIP_CHECK(hipGetLastError());
    s->om_level[l] = s->prm.omega;
    if (s->learn && l < int(s->learn->lv.size())) {
      s->learn->lv[l].ready = false;
      if (l < int(s->learn->on.size())) s->learn->on[l] = 0;
    }
    if (l == 0 || (l == 1 && s->prm.coarse_op == DGH_COARSE_JAC3)) coarse = true;
  }
  HIP_CHECK(hipStreamSynchronize(s->st));
  if (coarse) setup_coarse(s);
}

}  // namespace dgh

// ================================================================ extern "C"
extern "C" {

int dgh_init(int device) {
  return guarded([&] {
    if (device < 0 || device >= State::MAX_DEVICES) fail(DGH_ERR_ARG, "device %d out of range", device);
    HIP_CHECK(hipSetDevice(device));
    g.device = device;
    if (!g.noalloc_part[device]) {   // once per device, with that device current
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        fail(DGH_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      g.noalloc_part[device] = static_cast<double*>(p);
    }
    int lds = 0;
    HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    g.max_lds = lds > 0 ? lds : 65536;
    std::vector<double> L(9216), M(9216), W(144 * 27 * 8), K(NSTENCIL * KSTRIDE);
    generate_stencils(L.data(), M.data(), nullptr, nullptr);
    generate_transfer_weights(W.data());
    for (int t = 0; t < NSTENCIL; ++t)   // child-0 stencils: [t][L triangle, M triangle]
      for (int k = 0; k < 8; ++k)
        for (int j = k; j < 8; ++j) {
          K[t * KSTRIDE + tri(k, j)] = L[size_t(t) * 64 + k * 8 + j];
          K[t * KSTRIDE + 36 + tri(k, j)] = M[size_t(t) * 64 + k * 8 + j];
        }
    g.K64.upload(K.data(), K.size());
    g.W64.upload(W.data(), W.size());
    std::vector<float> Kf(K.begin(), K.end()), Wf(W.begin(), W.end());
    g.K32.upload(Kf.data(), Kf.size());
    g.W32.upload(Wf.data(), Wf.size());
    std::vector<double> KF(size_t(NSTENCIL) * KFSTRIDE);
    for (int t = 0; t < NSTENCIL; ++t) {   // child-0 full rows: [t][L rows, M rows]
      g.lrow[t] = g.mrow[t] = 0.0;
      for (int k = 0; k < 8; ++k) {
        double ls = 0.0, ms = 0.0;
        for (int j = 0; j < 8; ++j) {
          KF[size_t(t) * KFSTRIDE + k * 8 + j] = L[size_t(t) * 64 + k * 8 + j];
          KF[size_t(t) * KFSTRIDE + 64 + k * 8 + j] = M[size_t(t) * 64 + k * 8 + j];
          ls += std::fabs(L[size_t(t) * 64 + k * 8 + j]);
          ms += std::fabs(M[size_t(t) * 64 + k * 8 + j]);
        }
        g.lrow[t] = std::max(g.lrow[t], ls);
        g.mrow[t] = std::max(g.mrow[t], ms);
      }
    }
    g.KF64.upload(KF.data(), KF.size());
    g.hKF = KF;
    std::vector<float> KFf(KF.begin(), KF.end());
    g.KF32.upload(KFf.data(), KFf.size());
    g.ready = true;
  });
}

int dgh_alloc(size_t bytes, void** ptr) {
  return guarded([&] {
    require_init();
    if (!ptr) fail(DGH_ERR_ARG, "NULL argument");
    *ptr = nullptr;
    if (bytes && hipMalloc(ptr, bytes) != hipSuccess) {
      *ptr = nullptr;
      fail(DGH_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
  });
}

int dgh_free(void* ptr) {
  return guarded([&] {
    require_init();
    if (ptr) HIP_CHECK(hipFree(ptr));
  });
}

int dgh_copy_to_device(void* dst_dev, const void* src_host, size_t bytes) {
  return guarded([&] {
    require_init();
    if (bytes && (!dst_dev || !src_host)) fail(DGH_ERR_ARG, "NULL argument");
    if (bytes) HIP_CHECK(hipMemcpy(dst_dev, src_host, bytes, hipMemcpyHostToDevice));
  });
}

int dgh_copy_to_host(void* dst_host, const void* src_dev, size_t bytes) {
  return guarded([&] {
    require_init();
    if (bytes && (!dst_host || !src_dev)) fail(DGH_ERR_ARG, "NULL argument");
    if (bytes) HIP_CHECK(hipMemcpy(dst_host, src_dev, bytes, hipMemcpyDeviceToHost));
  });
}

int dgh_synchronize(void) {
  return guarded([&] {
    require_init();
    HIP_CHECK(hipDeviceSynchronize());
  });
}

int dgh_finalize(void) {
  return guarded([&] {
    g.K64.release();
    g.W64.release();
    g.K32.release();
    g.W32.release();
    g.KF64.release();
    g.KF32.release();
    if (g.blas) rocblas_destroy_handle(g.blas);
    g.blas = nullptr;
    g.ready = false;
  });
}

int dgh_mesh_create(uint32_t n_elem, uint32_t n_node, uint32_t max_depth, const uint32_t* nlist,
                    const uint8_t* child, const uint8_t* hn_mask, const uint8_t* lev, dgh_mesh* out) {
  return guarded([&] {
    require_init();
    if (!out || (n_elem && (!nlist || !child || !hn_mask || !lev))) fail(DGH_ERR_ARG, "NULL argument");
    if (max_depth < 2 || max_depth > 21) fail(DGH_ERR_ARG, "max_depth %u out of range", max_depth);
    if (n_elem == 0 || n_node == 0) fail(DGH_ERR_ARG, "empty mesh");
    auto m = std::make_unique<dgh_mesh_s>();
    m->E = n_elem;
    m->N = n_node;
    m->maxD = max_depth;
    m->hFac = 1.0 / double(1u << (max_depth - 1));
    m->h_cs.resize(n_elem);
    for (uint32_t e = 0; e < n_elem; ++e) {
      const int t = etype_of(child[e], hn_mask[e]);
      if (t < 0) fail(DGH_ERR_MESH, "element %u: hanging mask %u invalid for child %u", e, hn_mask[e], child[e]);
      if (lev[e] < 1 || lev[e] >= max_depth + 1) fail(DGH_ERR_MESH, "element %u: level %u out of range", e, lev[e]);
      m->h_cs[e] = uint8_t(18 * child[e] + t);
      for (int k = 0; k < 8; ++k)
        if (nlist[size_t(e) * 8 + k] >= n_node) fail(DGH_ERR_MESH, "element %u: node index out of range", e);
    }
    m->h_nl.assign(nlist, nlist + size_t(n_elem) * 8);
    m->h_child.assign(child, child + n_elem);
    m->h_hn.assign(hn_mask, hn_mask + n_elem);
    m->h_lev.assign(lev, lev + n_elem);
    m->nl.upload(m->h_nl.data(), m->h_nl.size());
    m->cs.upload(m->h_cs.data(), n_elem);
    // default tiling: larger tiles (fewer halo records) where the tiles still fill the device
    if (n_elem >= LARGE_MESH_ELEMS) {
      m->block = 1024;
      plan_mesh(m.get(), 8192);
    } else {
      plan_mesh(m.get(), 2048);
    }
    *out = m.release();
  });
}

int dgh_mesh_set_tile(dgh_mesh m, uint32_t tile_elems, uint32_t block) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (block != 0 && block != 256 && block != 512 && block != 1024) fail(DGH_ERR_ARG, "block must be 256, 512 or 1024");
    if (block) m->block = int(block);
    plan_mesh(m, tile_elems);
  });
}

int dgh_mesh_set_lane_stencils(dgh_mesh m, int on) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (on != DGH_ROWS_TYPE && on != DGH_ROWS_LANE && on != DGH_ROWS_CLASS)
      fail(DGH_ERR_ARG, "stencil rows: DGH_ROWS_TYPE, DGH_ROWS_LANE or DGH_ROWS_CLASS");
    if (on == DGH_ROWS_LANE && tile_lds_bytes<double>(m->block, m->op.plan.max_touch, true) > size_t(g.max_lds))
      fail(DGH_ERR_ARG, "the tiling leaves no LDS room for the stencil table");
    ++m->gen;   // captured graphs hold the kernel choice
    m->lane_rows = on;
    m->rows_auto = false;
  });
}

int dgh_mesh_lane_stencils(dgh_mesh m, int* on) {
  return guarded([&] {
    if (!m || !on) fail(DGH_ERR_ARG, "NULL argument");
    *on = m->lane_rows;
  });
}

int dgh_mesh_tile_info(dgh_mesh m, uint32_t* tile_elems, uint32_t* n_tiles, uint64_t* n_records,
                       uint64_t* n_node_entries, uint32_t* lds_bytes_f64) {
  return guarded([&] {
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (tile_elems) *tile_elems = m->op.plan.tile_elems;
    if (n_tiles) *n_tiles = uint32_t(m->op.plan.tiles.size());
    if (n_records) *n_records = m->op.plan.rec_elem.size();
    if (n_node_entries) *n_node_entries = m->op.plan.node_tab.size();
    if (lds_bytes_f64) *lds_bytes_f64 = uint32_t(op_lds_bytes<double>(m));
  });
}

int dgh_mesh_autotune(dgh_mesh m, int prec, uint32_t* tile_elems, uint32_t* block, float* us) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (prec != DGH_F64 && prec != DGH_F32) fail(DGH_ERR_ARG, "unknown precision %d", prec);
    const size_t bytes = (prec == DGH_F64 ? 8 : 4) * size_t(m->N);
    DevBuf<double> xb((bytes + 7) / 8), yb((bytes + 7) / 8);
    HIP_CHECK(hipMemset(xb.p, 0, bytes));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    uint32_t bt = 0, bb = 0;
    int bl = 0;
    // small levels need small tiles: enough workgroups to fill the device
    for (uint32_t T : {128u, 256u, 512u, 1024u, 2048u, 4096u, 8192u}) {
      for (int B : {256, 512, 1024}) {
        m->block = B;
        try {
          plan_mesh(m, T);
        } catch (const Error&) {
          continue;
        }
        if (m->op.plan.tile_elems != T) continue;   // reduced to fit LDS: tried already
        // stencil rows per wavefront type, per lane from LDS, per (type, level) class
        for (int L : {DGH_ROWS_TYPE, DGH_ROWS_LANE, DGH_ROWS_CLASS}) {
          if (L == DGH_ROWS_CLASS && !m->h_kappa.empty()) continue;   // uniform coefficients only
          m->lane_rows = L;
          if (op_lds_bytes<double>(m) > size_t(g.max_lds)) continue;
          auto run = [&] {
            if (prec == DGH_F64)
              launch_matvec<double>(m, xb.p, yb.p, nullptr);
            else
              launch_matvec<float>(m, reinterpret_cast<const float*>(xb.p), reinterpret_cast<float*>(yb.p), nullptr);
          };
          for (int i = 0; i < 3; ++i) run();
          HIP_CHECK(hipEventRecord(e0, nullptr));
          for (int i = 0; i < 10; ++i) run();
          HIP_CHECK(hipEventRecord(e1, nullptr));
          HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            bt = T;
            bb = uint32_t(B);
            bl = L;
          }
        }
      }
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    if (!bt) fail(DGH_ERR_STATE, "no tiling candidate fits");
    m->block = int(bb);
    m->lane_rows = bl;
    m->rows_auto = false;
    plan_mesh(m, bt);
    if (tile_elems) *tile_elems = bt;
    if (block) *block = bb;
    if (us) *us = best * 100.f;   // ms per 10 launches -> us per launch
  });
}

int dgh_mesh_set_coefficients(dgh_mesh m, const double* kappa, const double* mass, double lap_fac,
                              double mass_fac) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (!kappa && mass) fail(DGH_ERR_ARG, "mass without kappa");
    m->lap_fac = lap_fac;
    m->mass_fac = mass_fac;
    m->h_kappa.clear();
    m->h_mass.clear();
    if (kappa) m->h_kappa.assign(kappa, kappa + m->E);
    if (mass) m->h_mass.assign(mass, mass + m->E);
    upload_coefs(m, m->op);
    if (!m->mop.plan.tiles.empty()) upload_coefs(m, m->mop);
    ++m->coef_gen;
  });
}

int dgh_mesh_info(dgh_mesh m, uint32_t* n_elem, uint32_t* n_node) {
  return guarded([&] {
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    if (n_elem) *n_elem = m->E;
    if (n_node) *n_node = m->N;
  });
}

int dgh_mesh_destroy(dgh_mesh m) {
  delete m;
  return DGH_OK;
}

int dgh_matvec(dgh_mesh m, int prec, const void* x, void* y, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !x || !y) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (prec == DGH_F64)
      launch_matvec<double>(m, static_cast<const double*>(x), static_cast<double*>(y), st);
    else if (prec == DGH_F32)
      launch_matvec<float>(m, static_cast<const float*>(x), static_cast<float*>(y), st);
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_matvec_dot(dgh_mesh m, const double* x, double* y, double* xy, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !x || !y || !xy) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t nt = m->op.plan.tiles.size();
    if (m->dotp.n < std::max<size_t>(nt, 1)) m->dotp.alloc(std::max<size_t>(nt, 1));
    if (nt) launch_cg(m, x, y, m->dotp.p, st);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(RED2_THREADS), 0, st, int(nt), (const double*)m->dotp.p, xy);
    HIP_CHECK(hipGetLastError());
  });
}

int dgh_matvec_dot_noalloc(dgh_mesh m, const double* x, double* y, double* xy, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !x || !y || !xy) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    const size_t nt = m->op.plan.tiles.size();
    if (nt > size_t(DGH_NOALLOC_MAX_TILES))
      fail(DGH_ERR_ARG, "mesh has %zu operator tiles; dgh_matvec_dot_noalloc holds at most %d (use larger tiles "
           "or dgh_matvec_dot)", nt, DGH_NOALLOC_MAX_TILES);
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= State::MAX_DEVICES || !g.noalloc_part[dev])
      fail(DGH_ERR_STATE, "dgh_init has not been called on device %d", dev);
    double* part = g.noalloc_part[dev];
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (nt) launch_cg(m, x, y, part, st);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(RED2_THREADS), 0, st, int(nt), (const double*)part, xy);
    HIP_CHECK(hipGetLastError());
  });
}

int dgh_diagonal(dgh_mesh m, int prec, void* d, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !d) fail(DGH_ERR_ARG, "NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (prec == DGH_F64)
      launch_diag<double>(m, static_cast<double*>(d), st);
    else if (prec == DGH_F32)
      launch_diag<float>(m, static_cast<float*>(d), st);
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_mesh_set_boundary(dgh_mesh m, const uint8_t* bdy) {
  return guarded([&] {
    require_init();
    if (!m) fail(DGH_ERR_ARG, "NULL mesh");
    ++m->bdy_gen;
    ++m->gen;   // captured graphs hold the flags' buffer
    m->bdy.release();
    m->h_bdy.clear();
    if (!bdy || !m->N) return;
    m->bdy.upload(bdy, m->N);
    m->h_bdy.assign(bdy, bdy + m->N);
  });
}

int dgh_matvec_dirichlet(dgh_mesh m, int prec, const void* x, void* y, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !x || !y) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    if (m->N && !m->bdy.p) fail(DGH_ERR_STATE, "no boundary flags (dgh_mesh_set_boundary)");
    if (!m->N) return;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // one operator launch: the gather stages boundary x as 0, the store writes x on boundary rows
    auto run = [&](auto* xt, auto* yt) {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(xt)>>;
      launch_op<T, OP_DIRMV>(m, OpIO<T>{xt, yt, nullptr, nullptr, T(0), nullptr, m->bdy.p}, st);
    };
    if (prec == DGH_F64)
      run(static_cast<const double*>(x), static_cast<double*>(y));
    else if (prec == DGH_F32)
      run(static_cast<const float*>(x), static_cast<float*>(y));
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_diagonal_dirichlet(dgh_mesh m, int prec, void* d, void* stream) {
  return guarded([&] {
    require_init();
    if (!m || !d) fail(DGH_ERR_ARG, "NULL argument");
    if (m->N && !m->bdy.p) fail(DGH_ERR_STATE, "no boundary flags (dgh_mesh_set_boundary)");
    if (!m->N) return;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the diagonal's store writes 1 on boundary nodes
    if (prec == DGH_F64)
      launch_op<double, OP_DIAG>(m, OpIO<double>{nullptr, static_cast<double*>(d), nullptr, nullptr, 0.0, nullptr,
                                                 m->bdy.p}, st);
    else if (prec == DGH_F32)
      launch_op<float, OP_DIAG>(m, OpIO<float>{nullptr, static_cast<float*>(d), nullptr, nullptr, 0.f, nullptr,
                                               m->bdy.p}, st);
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_transfer_create(dgh_mesh coarse, dgh_mesh fine, const uint32_t* coarse_elem, const uint8_t* pos,
                        dgh_transfer* out) {
  return guarded([&] {
    require_init();
    if (!coarse || !fine || !out || (fine->N && (!coarse_elem || !pos))) fail(DGH_ERR_ARG, "NULL argument");
    if (coarse->maxD != fine->maxD) fail(DGH_ERR_ARG, "levels of different max depth");
    auto t = std::make_unique<dgh_transfer_s>();
    t->coarse = coarse;
    t->fine = fine;
    const uint32_t nf = fine->N, nc = coarse->N;
    for (uint32_t j = 0; j < nf; ++j) {
      if (coarse_elem[j] >= coarse->E) fail(DGH_ERR_MESH, "fine node %u: coarse element out of range", j);
      if (pos[j] >= 27) fail(DGH_ERR_MESH, "fine node %u: position code out of range", j);
    }
    t->celem.upload(coarse_elem, nf);
    t->pos.upload(pos, nf);
    t->h_celem.assign(coarse_elem, coarse_elem + nf);
    t->h_pos.assign(pos, pos + nf);
    // R = P^T with P[j][cnl[k]] = W[s][pos_j][k]: rows are coarse nodes, columns fine nodes
    std::vector<double> W(144 * 27 * 8);
    generate_transfer_weights(W.data());
    std::vector<uint32_t> rowptr(size_t(nc) + 1, 0);
    for (uint32_t j = 0; j < nf; ++j) {
      const uint32_t ce = coarse_elem[j];
      const double* w = W.data() + (size_t(coarse->h_cs[ce]) * 27 + pos[j]) * 8;
      for (int k = 0; k < 8; ++k)
        if (w[k] != 0.0) rowptr[coarse->h_nl[size_t(ce) * 8 + k] + 1]++;
    }
    for (uint32_t i = 0; i < nc; ++i) rowptr[i + 1] += rowptr[i];
    std::vector<uint32_t> col(rowptr[nc]), fillp(rowptr.begin(), rowptr.end() - 1);
    std::vector<float> wv(rowptr[nc]);
    for (uint32_t j = 0; j < nf; ++j) {
      const uint32_t ce = coarse_elem[j];
      const double* w = W.data() + (size_t(coarse->h_cs[ce]) * 27 + pos[j]) * 8;
      for (int k = 0; k < 8; ++k) {
        if (w[k] == 0.0) continue;
        const uint32_t q = fillp[coarse->h_nl[size_t(ce) * 8 + k]]++;
        col[q] = j;
        wv[q] = float(w[k]);
        if (double(wv[q]) != w[k]) fail(DGH_ERR_MESH, "transfer weight %.17g is not exact in FP32", w[k]);
      }
    }
    t->rowptr.upload(rowptr.data(), rowptr.size());
    t->col.upload(col.data(), col.size());
    t->w.upload(wv.data(), wv.size());
    // packed entries (half the index traffic of the restriction) when every weight is a
    // multiple of 1/32 in (0, 1] and the fine nodes fit in 27 bits
    bool packable = nf <= RP_COL + 1u;
    std::vector<uint32_t> ent(col.size());
    for (size_t q = 0; q < col.size() && packable; ++q) {
      const double k = double(wv[q]) * 32.0;
      packable = (k == std::floor(k) && k >= 1.0 && k <= 32.0);
      ent[q] = (uint32_t(k - 1.0) << RP_SHIFT) | col[q];
    }
    if (packable) t->ent.upload(ent.data(), ent.size());
    *out = t.release();
  });
}

int dgh_transfer_destroy(dgh_transfer t) {
  delete t;
  return DGH_OK;
}

int dgh_restrict(dgh_transfer t, int prec, const void* f, void* c, void* stream) {
  return guarded([&] {
    require_init();
    if (!t || !f || !c) fail(DGH_ERR_ARG, "NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (prec == DGH_F64)
      launch_restrict<double>(t, static_cast<const double*>(f), static_cast<double*>(c), st);
    else if (prec == DGH_F32)
      launch_restrict<float>(t, static_cast<const float*>(f), static_cast<float*>(c), st);
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_matvec_jac3(dgh_transfer t, int prec, const void* x, void* y, void* stream) {
  return guarded([&] {
    require_init();
    if (!t || !x || !y) fail(DGH_ERR_ARG, "NULL argument");
    if (x == y) fail(DGH_ERR_ARG, "x and y must not alias");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t nf = t->fine->N;
    if (prec == DGH_F64) {
      if (t->j3v64.n < nf) {
        t->j3v64.alloc(nf);
        t->j3w64.alloc(nf);
      }
      launch_jac3<double>(t, static_cast<const double*>(x), static_cast<double*>(y), t->j3v64.p, t->j3w64.p, st);
    } else if (prec == DGH_F32) {
      if (t->j3v32.n < nf) {
        t->j3v32.alloc(nf);
        t->j3w32.alloc(nf);
      }
      launch_jac3<float>(t, static_cast<const float*>(x), static_cast<float*>(y), t->j3v32.p, t->j3w32.p, st);
    } else {
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
    }
  });
}

int dgh_prolong_add(dgh_transfer t, int prec, const void* c, const void* f_in, void* f_out, void* stream) {
  return guarded([&] {
    require_init();
    if (!t || !c || !f_out) fail(DGH_ERR_ARG, "NULL argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (prec == DGH_F64)
      launch_prolong<double>(t, static_cast<const double*>(c), static_cast<const double*>(f_in),
                             static_cast<double*>(f_out), st);
    else if (prec == DGH_F32)
      launch_prolong<float>(t, static_cast<const float*>(c), static_cast<const float*>(f_in),
                            static_cast<float*>(f_out), st);
    else
      fail(DGH_ERR_ARG, "unknown precision %d", prec);
  });
}

int dgh_solver_default_params(dgh_solver_params* p) {
  if (!p) return DGH_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->rtol = 1e-8;         // -ksp_rtol
  p->omega = 0.667;       // -mg_levels_ksp_richardson_scale
  p->max_it = 10;         // -ksp_max_it
  p->nu_pre = 4;          // -mg_levels_ksp_max_it
  p->nu_post = 4;
  p->norm_type = DGH_NORM_PRECONDITIONED;
  p->null_space = 0;
  p->vcycle_prec = DGH_F64;
  p->smoother = DGH_SMOOTHER_JACOBI;
  p->use_graph = 0;
  p->coarse_op = DGH_COARSE_JAC2;
  p->coarse_dense_max = COARSE_DENSE_LIMIT;
  p->coarse_rtol = 1e-12;
  p->coarse_max_it = 5000;
  return DGH_OK;
}

int dgh_solver_create(int nlevels, const dgh_mesh* meshes, const dgh_transfer* transfers,
                      const dgh_solver_params* params, dgh_solver* out) {
  return guarded([&] {
    require_init();
    if (nlevels < 1 || !meshes || !params || !out || (nlevels > 1 && !transfers))
      fail(DGH_ERR_ARG, "invalid arguments");
    if (params->nu_pre < 1 || params->nu_post < 0 || params->max_it < 0 || !(params->omega > 0))
      fail(DGH_ERR_ARG, "invalid smoother parameters");
    if (params->vcycle_prec != DGH_F64 && params->vcycle_prec != DGH_F32) fail(DGH_ERR_ARG, "bad vcycle_prec");
    if (params->smoother != DGH_SMOOTHER_JACOBI && params->smoother != DGH_SMOOTHER_LEARNED)
      fail(DGH_ERR_ARG, "unknown smoother %d", params->smoother);
    if (params->coarse_op != DGH_COARSE_JAC2 && params->coarse_op != DGH_COARSE_JAC3)
      fail(DGH_ERR_ARG, "unknown coarse operator %d", params->coarse_op);
    if (params->coarse_dense_max < 0 || !(params->coarse_rtol > 0) || params->coarse_max_it < 1)
      fail(DGH_ERR_ARG, "invalid coarse solve parameters");
    if (params->coarse_op == DGH_COARSE_JAC3 && params->smoother == DGH_SMOOTHER_LEARNED)
      fail(DGH_ERR_ARG, "the learned smoother is trained on Jacobian2 levels; use DGH_COARSE_JAC2");
    auto s = std::make_unique<dgh_solver_s>();
    s->nlevels = nlevels;
    s->prm = *params;
    s->om_level.assign(nlevels, params->omega);
    s->meshes.assign(meshes, meshes + nlevels);
    if (nlevels > 1) s->transfers.assign(transfers, transfers + nlevels - 1);
    for (int l = 0; l < nlevels; ++l) {
      if (!meshes[l]) fail(DGH_ERR_ARG, "NULL mesh at level %d", l);
      if (l > 0 && (s->transfers[l - 1]->coarse != meshes[l - 1] || s->transfers[l - 1]->fine != meshes[l]))
        fail(DGH_ERR_ARG, "transfer %d does not connect levels %d and %d", l - 1, l - 1, l);
    }
    // boundary flags on every level: the Dirichlet problem (SetDirichletJacContexts,
    // examples/src/backend/omgJac.C:508-586); on some levels only: an error
    int nb = 0;
    for (int l = 0; l < nlevels; ++l) nb += (meshes[l]->N == 0 || meshes[l]->bdy.p) ? 1 : 0;
    if (nb != 0 && nb != nlevels)
      fail(DGH_ERR_STATE, "boundary flags are set on %d of %d levels (all or none)", nb, nlevels);
    s->dirichlet = (nb == nlevels) && meshes[nlevels - 1]->bdy.p != nullptr;
    if (s->dirichlet && (params->null_space || params->smoother == DGH_SMOOTHER_LEARNED ||
                         params->coarse_op != DGH_COARSE_JAC2))
      fail(DGH_ERR_ARG, "the Dirichlet solve takes no null space, learned smoother or Jacobian3 levels");
    for (int l = 0; l < nlevels; ++l) {
      s->coef_gen.push_back(meshes[l]->coef_gen);
      s->bdy_gen.push_back(meshes[l]->bdy_gen);
    }
    HIP_CHECK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s->hS), sizeof(double) * S_NSCAL, hipHostMallocDefault));
    s->S.alloc(S_NSCAL);
    s->part.alloc(size_t(std::max(MAXQ, NDOT)) * RED_BLOCKS);
    s->part_p.alloc(MAXQ * RED_BLOCKS);
    // per-level vectors and Jacobi diagonals
    s->v64.resize(nlevels);
    if (params->vcycle_prec == DGH_F32) s->v32.resize(nlevels);
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = meshes[l]->N;
      auto& V = s->v64[l];
      V.invd.alloc(n);
      V.b.alloc(n);
      V.xa.alloc(n);
      V.xb.alloc(n);
      V.r.alloc(n);
      const bool j3fine = params->coarse_op == DGH_COARSE_JAC3 && l >= 1;   // fine level of a Jacobian3 level
      if (j3fine) V.t.alloc(n);
      solver_diag(s.get(), l, V.r.p);
      hipLaunchKernelGGL(reciprocal_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, s->st, n,
                         (const double*)V.r.p, V.invd.p);
      if (params->vcycle_prec == DGH_F32) {
        auto& F = s->v32[l];
        F.invd.alloc(n);
        F.b.alloc(n);
        F.xa.alloc(n);
        F.xb.alloc(n);
        F.r.alloc(n);
        if (j3fine) F.t.alloc(n);
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n)), dim3(VBLOCK), 0, s->st, n,
                           (const double*)V.invd.p, F.invd.p);
      }
    }
    HIP_CHECK(hipStreamSynchronize(s->st));
    setup_coarse(s.get());
    const uint32_t nf = meshes[nlevels - 1]->N;
    s->x.alloc(nf);
    s->r.alloc(nf);
    s->z.alloc(nf);
    s->p.alloc(nf);
    s->w.alloc(nf);
    *out = s.release();
  });
}

int dgh_solver_destroy(dgh_solver s) {
  delete s;
  return DGH_OK;
}

int dgh_solver_coarse_info(dgh_solver s, int* iterative, int* last_iterations) {
  return guarded([&] {
    if (!s) fail(DGH_ERR_ARG, "NULL solver");
    if (iterative) *iterative = s->coarse_iter ? 1 : 0;
    if (last_iterations) *last_iterations = s->coarse_its;
  });
}

int dgh_solver_vcycle(dgh_solver s, const double* r, double* z, void* stream) {
  return guarded([&] {
    require_init();
    if (!s || !r || !z) fail(DGH_ERR_ARG, "NULL argument");
    refresh_operators(s);
    precondition(s, r, z, static_cast<hipStream_t>(stream));
  });
}

// PETSc KSPSolve_CG restated (zero initial guess, left preconditioning): see the oracle's
// or_pcg.  In the FP32-V-cycle mode the preconditioner is inexact and the p-update uses the
// flexible (Polak-Ribiere) factor z_k.(r_k - r_{k-1}) / z_{k-1}.r_{k-1}.  CG scalars live on
// the device; the host reads the stopping-test norm once per iteration.
int dgh_solver_solve(dgh_solver s, const double* b, double* x, dgh_solve_stats* stats, double* history,
                     int history_len, void* stream) {
  return guarded([&] {
    require_init();
    if (!s || !b || !x) fail(DGH_ERR_ARG, "NULL argument");
    hipStream_t ust = static_cast<hipStream_t>(stream);
    hipStream_t st = s->st;
    const int top = s->nlevels - 1;
    dgh_mesh m = s->meshes[top];
    const uint32_t n = m->N;
    const bool pnorm = (s->prm.norm_type == DGH_NORM_PRECONDITIONED);
    double* S = s->S.p;
    hipEvent_t e0, e1, eu;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventCreateWithFlags(&eu, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(eu, ust));          // b is ready on the caller's stream
    HIP_CHECK(hipStreamWaitEvent(st, eu, 0));
    HIP_CHECK(hipEventRecord(e0, st));
    const bool mixed = (s->prm.vcycle_prec == DGH_F32);
    refresh_operators(s);
    const size_t ntiles = std::max<size_t>(m->op.plan.tiles.size(), size_t(RED_BLOCKS) * MAXQ);
    if (s->ptile.n < ntiles) s->ptile.alloc(ntiles);
    const uint64_t gen = mesh_generation(s);
    if (gen != s->graph_gen) {   // re-tiled or new coefficients since the graphs were captured
      drop_graphs(s);
      s->graph_gen = gen;
    }
    // r = b (projected), x = 0, p = 0, z = B r, dots
    hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, b, s->r.p);
    if (s->prm.null_space) project_mean<double>(s, n, s->r.p, st);
    hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, s->x.p, 0.0, n);
    hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, s->p.p, 0.0, n);
    if (mixed)
      hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n,
                         (const double*)s->r.p, s->v32[top].b.p);
    s->zres = precondition_raw(s, s->r.p, st);
    if (mixed)
      hipLaunchKernelGGL(cg_dots_kernel<float>, dim3(DOT_BLOCKS), dim3(VBLOCK), 0, st, n,
                         static_cast<const float*>(s->zres), (const double*)s->r.p, (const double*)nullptr, s->part.p);
    else
      hipLaunchKernelGGL(cg_dots_kernel<double>, dim3(DOT_BLOCKS), dim3(VBLOCK), 0, st, n,
                         static_cast<const double*>(s->zres), (const double*)s->r.p, (const double*)nullptr, s->part.p);
    hipLaunchKernelGGL(cg_init_fused_kernel, dim3(1), dim3(RED2_THREADS), 0, st, (const double*)s->part.p, n,
                       int(s->prm.null_space), S, int(pnorm));
    HIP_CHECK(hipMemcpyAsync(s->hS, S, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const double dp0 = s->hS[S_DP];
    const double bnorm = std::sqrt(s->hS[S_BB]);
    double dp = dp0;
    if (history && history_len > 0) history[0] = dp0;
    if (s->prm.use_graph && !s->gexec && !s->coarse_iter) {   // the iterative coarse solve reads back
      HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      cg_iteration(s, st);
      HIP_CHECK(hipStreamEndCapture(st, &s->graph));
      HIP_CHECK(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
    }
    int it = 0;
    bool conv = (dp0 == 0.0), bad = false;
    while (!conv && it < s->prm.max_it) {
      if (s->gexec)
        HIP_CHECK(hipGraphLaunch(s->gexec, st));
      else
        cg_iteration(s, st);
      HIP_CHECK(hipMemcpyAsync(&s->hS[S_DP], S + S_DP, sizeof(double), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      dp = s->hS[S_DP];
      ++it;
      if (history && it < history_len) history[it] = dp;
      if (!std::isfinite(dp)) {
        bad = true;
        break;
      }
      if (dp <= s->prm.rtol * dp0) conv = true;
    }
    hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, (const double*)s->x.p, x);
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (stats) {
      // true residual ||b - A x|| / ||b|| in FP64 (b projected when constants are the null space)
      hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n, b, s->w.p);
      if (s->prm.null_space) project_mean<double>(s, n, s->w.p, st);
      if (s->dirichlet) {
        launch_dirmv<double>(m, x, s->z.p, st);
        hipLaunchKernelGGL(resid_vec_kernel<double>, dim3(blocks_for(n)), dim3(VBLOCK), 0, st, n,
                           (const double*)s->w.p, (const double*)s->z.p, s->z.p);
      } else {
        launch_resid<double>(m, x, s->w.p, s->z.p, st);
      }
      DotArgs<double> q{};
      q.a[0] = s->z.p; q.b[0] = s->z.p;
      reduce<double>(n, q, 1, s->part_p.p, S + S_PROJ, st);
      HIP_CHECK(hipMemcpyAsync(s->hS, S, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      stats->iterations = it;
      stats->converged = conv ? 1 : 0;
      stats->rnorm0 = dp0;
      stats->rnorm = dp;
      stats->rel_true = bnorm > 0 ? std::sqrt(s->hS[S_PROJ]) / bnorm : 0.0;
      stats->seconds = ms * 1e-3;
    }
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    HIP_CHECK(hipEventDestroy(eu));
    if (bad) fail(DGH_ERR_DIVERGED, "CG breakdown: non-finite residual at iteration %d", it);
  });
}

}  // extern "C"
