// Learned Jacobi damping behind the C ABI.
//
// A per-element MLP maps element features (level, hanging type, child number, the 8 nodal
// log-coefficient ratios) to 8 damping logits, one per vertex; the logits are averaged per
// node and turned into weights omega_n = cap_n sigmoid(theta_n), where cap_n keeps every
// weighted sweep an A-norm contraction.  The smoother then runs x += omega_n / A_nn (b - A x)
// on that level.  Training data come from the solver: errors left after the pre-smoothing
// and coarse-grid correction of the V-cycle, on which the objective is the A-norm reduction
// of nu damped sweeps.  The network is trained on the device by the forward/backward kernels
// of learn_kernels.hpp (BF16 weights, FP32 accumulation, Adam).  A level keeps the learned
// weights only when they beat plain damped Jacobi on fresh samples of the same objective;
// otherwise it falls back to Jacobi.  There is no counterpart in the reference, whose
// smoother is damped Jacobi (bin/options:187-193).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "learn_kernels.hpp"

namespace dgh {

constexpr int LGRID = 512;     // workgroups of the MLP kernels

static LearnState& lstate(dgh_solver s) {
  if (!s->learn) s->learn = new LearnState();
  LearnState& L = *s->learn;
  if (L.lv.size() != size_t(s->nlevels)) {
    L.lv.clear();
    L.lv.resize(s->nlevels);
    L.on.assign(s->nlevels, 0);
  }
  return L;
}

static uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static void init_network(dgh_solver s, uint64_t seed) {
  LearnState& L = lstate(s);
  std::vector<float> th(LNPARAM, 0.f);
  uint64_t st = seed;
  auto normal = [&]() {
    const double u1 = (double(splitmix(st) >> 11) + 0.5) / 9007199254740992.0;
    const double u2 = double(splitmix(st) >> 11) / 9007199254740992.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int i = 0; i < LNH * LNF; ++i) th[L_W1 + i] = float(normal() * std::sqrt(2.0 / LNF));
  for (int i = 0; i < LNH * LNH; ++i) th[L_W2 + i] = float(normal() * std::sqrt(2.0 / LNH));
  for (int i = 0; i < LNO * LNH; ++i) th[L_W3 + i] = float(normal() * 0.01);
  // start at the configured damping where the cap is 1
  const double w0 = std::min(std::max(s->prm.omega, 0.05), 0.95);
  for (int m = 0; m < LNO; ++m) th[L_B3 + m] = float(std::log(w0 / (1.0 - w0)));
  L.theta.upload(th.data(), LNPARAM);
  std::vector<float> zero(LNPARAM, 0.f);
  L.m1.upload(zero.data(), LNPARAM);
  L.m2.upload(zero.data(), LNPARAM);
  L.grad.alloc(LNPARAM);
  L.gpart.alloc(size_t(LGRID) * LNPARAM);
  L.wbf.alloc(LNPARAM);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(LNPARAM)), dim3(256), 0, s->st, LNPARAM, (const float*)L.theta.p,
                     L.wbf.p);
  HIP_CHECK(hipGetLastError());
  L.dots.alloc(2 * MAXK);
  L.step = 0;
  L.init = true;
}

static LearnState& ready_state(dgh_solver s) {
  LearnState& L = lstate(s);
  if (!L.init) init_network(s, 0x5eedull);
  return L;
}

// Element features, node CSR, caps and plain inverse diagonals of one level (host once).
static void setup_level(dgh_solver s, int l) {
  LearnState& L = ready_state(s);
  LearnLevel& v = L.lv[l];
  if (v.ready) return;
  dgh_mesh m = s->meshes[l];
  const uint32_t E = m->E, N = m->N;
  std::vector<double> ksum(N, 0.0), d(N, 0.0), r(N, 0.0);
  std::vector<uint32_t> rowptr(size_t(N) + 1, 0);
  std::vector<float> lke(E), lkbar(N), cap(N);
  for (uint32_t e = 0; e < E; ++e) {
    const double k = m->h_kappa.empty() ? m->lap_fac : m->h_kappa[e];
    lke[e] = float(std::log10(k));
    for (int q = 0; q < 8; ++q) {
      ksum[m->h_nl[size_t(e) * 8 + q]] += k;
      rowptr[m->h_nl[size_t(e) * 8 + q] + 1]++;
    }
  }
  for (uint32_t n = 0; n < N; ++n) {
    lkbar[n] = rowptr[n + 1] ? float(std::log10(ksum[n] / rowptr[n + 1])) : 0.f;
    rowptr[n + 1] += rowptr[n];
  }
  std::vector<uint32_t> slot(size_t(E) * 8), fillp(rowptr.begin(), rowptr.end() - 1);
  for (uint32_t e = 0; e < E; ++e)
    for (int q = 0; q < 8; ++q) slot[fillp[m->h_nl[size_t(e) * 8 + q]]++] = e * 8 + q;
  // cap_n = 2 A_nn / sum |K_e[k][j]| over the element rows at n
  std::vector<double> L2(9216), M2(9216);
  generate_stencils(L2.data(), M2.data(), nullptr, nullptr);
  for (uint32_t e = 0; e < E; ++e) {
    const double h = m->hFac * double(1u << (m->maxD - m->h_lev[e]));
    const double k1 = m->h_kappa.empty() ? m->lap_fac : m->h_kappa[e];
    const double k2 = m->h_kappa.empty() ? m->mass_fac : (m->h_mass.empty() ? 0.0 : m->h_mass[e]);
    const double f1 = k1 * h / 2.0, f2 = k2 * h * h * h / 8.0;
    const double* Ls = L2.data() + size_t(m->h_cs[e]) * 64;
    const double* Ms = M2.data() + size_t(m->h_cs[e]) * 64;
    const uint32_t* idx = m->h_nl.data() + size_t(e) * 8;
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 8; ++j) {
        const double a = f1 * Ls[k * 8 + j] + f2 * Ms[k * 8 + j];
        r[idx[k]] += std::fabs(a);
        if (idx[j] == idx[k]) d[idx[k]] += a;
      }
  }
  for (uint32_t n = 0; n < N; ++n) cap[n] = r[n] > 0 ? float(2.0 * d[n] / r[n]) : 1.f;
  v.lev.upload(m->h_lev.data(), E);
  v.lke.upload(lke.data(), E);
  v.lkbar.upload(lkbar.data(), N);
  v.cap.upload(cap.data(), N);
  v.rowptr.upload(rowptr.data(), rowptr.size());
  v.slot.upload(slot.data(), slot.size());
  // plain Jacobi inverse diagonals (the level's smoother vectors hold them until weights
  // are installed)
  v.dinv64.alloc(N);
  HIP_CHECK(hipMemcpyAsync(v.dinv64.p, s->v64[l].invd.p, sizeof(double) * N, hipMemcpyDeviceToDevice, s->st));
  v.dinv32.alloc(N);
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(N)), dim3(VBLOCK), 0, s->st, N,
                     (const double*)v.dinv64.p, v.dinv32.p);
  v.O.alloc(size_t(E) * 8);
  v.dO.alloc(size_t(E) * 8);
  v.theta_n.alloc(N);
  v.omega.alloc(N);
  v.gomega.alloc(N);
  HIP_CHECK(hipStreamSynchronize(s->st));
  v.ready = true;
}

static LearnElems elems(dgh_solver s, int l) {
  dgh_mesh m = s->meshes[l];
  LearnLevel& v = s->learn->lv[l];
  return LearnElems{m->nl.p, m->cs.p, v.lev.p, v.lke.p, v.lkbar.p, m->E, m->maxD};
}

// O = MLP(level l), then theta_n and omega
static void forward(dgh_solver s, int l) {
  LearnState& L = *s->learn;
  LearnLevel& v = L.lv[l];
  const uint32_t ntiles = (s->meshes[l]->E + LTE - 1) / LTE;
  const unsigned grid = std::min<uint32_t>(ntiles, LGRID);
  hipLaunchKernelGGL(mlp_kernel<false>, dim3(grid), dim3(LBLOCK), 0, s->st, elems(s, l), (const uint16_t*)L.wbf.p,
                     (const float*)L.theta.p, v.O.p, (const float*)nullptr, (float*)nullptr);
  HIP_CHECK(hipGetLastError());
  const uint32_t N = s->meshes[l]->N;
  hipLaunchKernelGGL(nodal_kernel, dim3(blocks_for(N)), dim3(256), 0, s->st, N, (const uint32_t*)v.rowptr.p,
                     (const uint32_t*)v.slot.p, (const float*)v.O.p, (const float*)v.cap.p, (const float*)v.dinv32.p,
                     v.theta_n.p, v.omega.p, (float*)nullptr, (double*)nullptr, (const double*)nullptr);
  HIP_CHECK(hipGetLastError());
}

// grad = d(sum dO . O)/d(params) over level l
static void backward(dgh_solver s, int l) {
  LearnState& L = *s->learn;
  LearnLevel& v = L.lv[l];
  const uint32_t ntiles = (s->meshes[l]->E + LTE - 1) / LTE;
  const unsigned grid = std::min<uint32_t>(ntiles, LGRID);
  hipLaunchKernelGGL(mlp_kernel<true>, dim3(grid), dim3(LBLOCK), 0, s->st, elems(s, l), (const uint16_t*)L.wbf.p,
                     (const float*)L.theta.p, (float*)nullptr, (const float*)v.dO.p, L.gpart.p);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(grad_reduce_kernel, dim3(blocks_for(LNPARAM)), dim3(256), 0, s->st, int(grid),
                     (const float*)L.gpart.p, L.grad.p);
  HIP_CHECK(hipGetLastError());
}

static void scratch(dgh_solver s, int S, int nu) {
  LearnState& L = *s->learn;
  const size_t ne = size_t(s->meshes[s->nlevels - 1]->N) * S;
  if (L.t.n < ne) {
    L.t.alloc(ne);
    L.g.alloc(ne);
    L.u.alloc(ne);
  }
  if (L.es.n < ne * nu) L.es.alloc(ne * nu);
  if (L.z.n < ne * nu) L.z.alloc(ne * nu);
}

// Objective of level l on samples e0 [N][S] for weights omega: returns the loss; with grad,
// v.gomega = dL/domega.
static double objective(dgh_solver s, int l, const float* omega, const float* e0, int S, int nu, bool grad,
                        float* gomega) {
  LearnState& L = *s->learn;
  LearnLevel& v = L.lv[l];
  dgh_mesh m = s->meshes[l];
  const uint32_t N = m->N;
  const size_t ne = size_t(N) * S;
  hipStream_t st = s->st;
  scratch(s, S, nu);
  ensure_multi(m);
  auto es = [&](int k) -> float* { return k == 0 ? const_cast<float*>(e0) : L.es.p + size_t(k - 1) * ne; };
  auto zk = [&](int k) { return L.z.p + size_t(k) * ne; };
  multi_op_f32(m, OP_MATVEC, e0, L.t.p, nullptr, nullptr, 0.f, S, st);
  column_dots_f32(N, S, e0, L.t.p, s->mpart.p, L.dots.p, st);
  for (int k = 0; k < nu; ++k) {
    if (k > 0) multi_op_f32(m, OP_MATVEC, es(k), L.t.p, nullptr, nullptr, 0.f, S, st);
    hipLaunchKernelGGL(lsweep_kernel, dim3(blocks_for(ne)), dim3(256), 0, st, N, S, (const float*)v.dinv32.p, omega,
                       (const float*)L.t.p, (const float*)es(k), zk(k), es(k + 1));
  }
  multi_op_f32(m, OP_MATVEC, es(nu), L.t.p, nullptr, nullptr, 0.f, S, st);
  column_dots_f32(N, S, es(nu), L.t.p, s->mpart.p, L.dots.p + MAXK, st);
  if (grad) {
    hipLaunchKernelGGL(lgrad_init_kernel, dim3(blocks_for(ne)), dim3(256), 0, st, N, S, (const float*)L.t.p,
                       (const double*)L.dots.p, L.g.p);
    for (int k = nu - 1; k >= 0; --k) {
      hipLaunchKernelGGL(lgrad_step_kernel, dim3(blocks_for(N)), dim3(256), 0, st, N, S, int(k != nu - 1),
                         (const float*)L.g.p, (const float*)zk(k), (const float*)v.dinv32.p, omega, gomega, L.u.p);
      if (k > 0) {
        multi_op_f32(m, OP_MATVEC, L.u.p, L.t.p, nullptr, nullptr, 0.f, S, st);
        hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(ne)), dim3(256), 0, st, ne, L.g.p, (const float*)L.t.p);
      }
    }
  }
  HIP_CHECK(hipGetLastError());
  double h[2 * MAXK];
  HIP_CHECK(hipMemcpyAsync(h, L.dots.p, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  double loss = 0.0;
  for (int j = 0; j < S; ++j) loss += h[MAXK + j] / h[j];
  return loss / S;
}

// Errors the post-smoother of level l faces: e = v - (pre-smoothing + coarse correction of
// A v) for S random v, with the smoothers currently installed.
static void gen_samples(dgh_solver s, int l, int S, uint64_t seed) {
  LearnState& L = *s->learn;
  LearnLevel& v = L.lv[l];
  ensure_multi_buffers(s, S);
  dgh_mesh m = s->meshes[l];
  const uint32_t N = m->N;
  const size_t ne = size_t(N) * S;
  hipStream_t st = s->st;
  if (v.e0.n < ne) v.e0.alloc(ne);
  auto& V = s->m32[l];
  const float om = float(s->om_level[l]);
  const float* invd = s->v32[l].invd.p;
  hipLaunchKernelGGL(rand_kernel, dim3(blocks_for(ne)), dim3(256), 0, st, ne, seed, v.e0.p);
  if (s->prm.null_space) project_mean_multi_f32(s, N, S, v.e0.p, st);
  multi_op_f32(m, OP_MATVEC, v.e0.p, V.b.p, nullptr, nullptr, 0.f, S, st);
  float* xc = V.xa.p;
  float* xn = V.xb.p;
  hipLaunchKernelGGL(smooth_first_multi_kernel<float>, dim3(blocks_for(ne)), dim3(VBLOCK), 0, st, N, S, om,
                     (const float*)V.b.p, invd, xc);
  for (int it = 1; it < s->prm.nu_pre; ++it) {
    multi_op_f32(m, OP_SMOOTH, xc, xn, V.b.p, invd, om, S, st);
    std::swap(xc, xn);
  }
  multi_op_f32(m, OP_RESID, xc, V.r.p, V.b.p, nullptr, 0.f, S, st);
  restrict_multi_f32(s->transfers[l - 1], S, V.r.p, s->m32[l - 1].b.p, st);
  const float* xcoarse = mvcycle(s, l - 1, S, s->m32[l - 1].b.p, st);
  prolong_multi_f32(s->transfers[l - 1], S, xcoarse, xc, xc, st);
  hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(ne)), dim3(256), 0, st, ne, v.e0.p, (const float*)xc);
  if (s->prm.null_space) project_mean_multi_f32(s, N, S, v.e0.p, st);
  HIP_CHECK(hipGetLastError());
  v.S = S;
  v.age = 0;
}

static void check_level(dgh_solver s, int l) {
  if (l < 1 || l >= s->nlevels) fail(DGH_ERR_ARG, "level %d has no smoother (levels 1..%d)", l, s->nlevels - 1);
}

static void check_samples(int S) {
  if (S < KS || S > MAXK || S % KS) fail(DGH_ERR_ARG, "samples %d: must be a multiple of %d in [%d, %d]", S, KS, KS, MAXK);
}

// install (on) or remove the learned weights of level l in the smoother vectors
static void install(dgh_solver s, int l, bool on) {
  LearnState& L = *s->learn;
  LearnLevel& v = L.lv[l];
  const uint32_t N = s->meshes[l]->N;
  if (on) {
    hipLaunchKernelGGL(nodal_kernel, dim3(blocks_for(N)), dim3(256), 0, s->st, N, (const uint32_t*)v.rowptr.p,
                       (const uint32_t*)v.slot.p, (const float*)v.O.p, (const float*)v.cap.p,
                       (const float*)v.dinv32.p, v.theta_n.p, v.omega.p,
                       s->v32.size() ? s->v32[l].invd.p : (float*)nullptr, s->v64[l].invd.p,
                       (const double*)v.dinv64.p);
    s->om_level[l] = 1.0;
  } else {
    HIP_CHECK(hipMemcpyAsync(s->v64[l].invd.p, v.dinv64.p, sizeof(double) * N, hipMemcpyDeviceToDevice, s->st));
    if (s->v32.size())
      HIP_CHECK(hipMemcpyAsync(s->v32[l].invd.p, v.dinv32.p, sizeof(float) * N, hipMemcpyDeviceToDevice, s->st));
    s->om_level[l] = s->prm.omega;
  }
  HIP_CHECK(hipGetLastError());
  L.on[l] = on ? 1 : 0;
}

}  // namespace dgh

extern "C" {

int dgh_learn_default_params(dgh_learn_params* p) {
  if (!p) return DGH_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->steps = 200;
  p->samples = 8;
  p->nu = 2;
  p->refresh = 4;
  p->lr = 3e-3f;
  p->seed = 5000;
  return DGH_OK;
}

int dgh_learned_init(dgh_solver s, uint64_t seed) {
  return guarded([&] {
    require_init();
    if (!s) fail(DGH_ERR_ARG, "NULL solver");
    init_network(s, seed);
    HIP_CHECK(hipStreamSynchronize(s->st));
  });
}

int dgh_learned_get_params(dgh_solver s, float* theta, int n) {
  return guarded([&] {
    require_init();
    if (!s || !theta) fail(DGH_ERR_ARG, "NULL argument");
    if (n != LNPARAM) fail(DGH_ERR_ARG, "the network has %d parameters", LNPARAM);
    LearnState& L = ready_state(s);
    HIP_CHECK(hipMemcpyAsync(theta, L.theta.p, sizeof(float) * LNPARAM, hipMemcpyDeviceToHost, s->st));
    HIP_CHECK(hipStreamSynchronize(s->st));
  });
}

int dgh_learned_set_params(dgh_solver s, const float* theta, int n) {
  return guarded([&] {
    require_init();
    if (!s || !theta) fail(DGH_ERR_ARG, "NULL argument");
    if (n != LNPARAM) fail(DGH_ERR_ARG, "the network has %d parameters", LNPARAM);
    LearnState& L = ready_state(s);
    HIP_CHECK(hipMemcpyAsync(L.theta.p, theta, sizeof(float) * LNPARAM, hipMemcpyHostToDevice, s->st));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(LNPARAM)), dim3(256), 0, s->st, LNPARAM,
                       (const float*)L.theta.p, L.wbf.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->st));
  });
}

int dgh_learned_eval(dgh_solver s, int level, const float* dO, float* O, float* grad) {
  return guarded([&] {
    require_init();
    if (!s) fail(DGH_ERR_ARG, "NULL solver");
    check_level(s, level);
    setup_level(s, level);
    LearnLevel& v = s->learn->lv[level];
    const size_t e8 = size_t(s->meshes[level]->E) * 8;
    forward(s, level);
    if (O) HIP_CHECK(hipMemcpyAsync(O, v.O.p, sizeof(float) * e8, hipMemcpyDeviceToHost, s->st));
    if (dO) {
      HIP_CHECK(hipMemcpyAsync(v.dO.p, dO, sizeof(float) * e8, hipMemcpyHostToDevice, s->st));
      backward(s, level);
      if (grad) HIP_CHECK(hipMemcpyAsync(grad, s->learn->grad.p, sizeof(float) * LNPARAM, hipMemcpyDeviceToHost, s->st));
    }
    HIP_CHECK(hipStreamSynchronize(s->st));
  });
}

int dgh_learned_loss(dgh_solver s, int level, const float* omega, const float* samples, int S, int nu,
                     uint64_t seed, double* loss, float* g_omega) {
  return guarded([&] {
    require_init();
    if (!s || !loss) fail(DGH_ERR_ARG, "NULL argument");
    check_level(s, level);
    check_samples(S);
    if (nu < 1 || nu > 8) fail(DGH_ERR_ARG, "nu %d out of range", nu);
    setup_level(s, level);
    ensure_multi_buffers(s, S);
    LearnLevel& v = s->learn->lv[level];
    const uint32_t N = s->meshes[level]->N;
    if (omega)
      HIP_CHECK(hipMemcpyAsync(v.omega.p, omega, sizeof(float) * N, hipMemcpyHostToDevice, s->st));
    else
      forward(s, level);
    if (samples) {
      if (v.e0.n < size_t(N) * S) v.e0.alloc(size_t(N) * S);
      HIP_CHECK(hipMemcpyAsync(v.e0.p, samples, sizeof(float) * N * S, hipMemcpyHostToDevice, s->st));
      v.S = S;
    } else {
      gen_samples(s, level, S, seed);
    }
    *loss = objective(s, level, v.omega.p, v.e0.p, S, nu, g_omega != nullptr, v.gomega.p);
    if (g_omega) {
      HIP_CHECK(hipMemcpyAsync(g_omega, v.gomega.p, sizeof(float) * N, hipMemcpyDeviceToHost, s->st));
      HIP_CHECK(hipStreamSynchronize(s->st));
    }
  });
}

int dgh_learned_train(dgh_solver s, const dgh_learn_params* p, dgh_learn_stats* stats) {
  return guarded([&] {
    require_init();
    if (!s || !p) fail(DGH_ERR_ARG, "NULL argument");
    if (s->nlevels < 2) fail(DGH_ERR_STATE, "a single level has no smoother");
    if (s->prm.coarse_op != DGH_COARSE_JAC2) fail(DGH_ERR_ARG, "learned damping is trained on Jacobian2 levels only");
    if (s->dirichlet) fail(DGH_ERR_ARG, "learned damping has no Dirichlet mode");
    refresh_operators(s);
    check_samples(p->samples);
    if (p->steps < 0 || p->nu < 1 || p->nu > 8 || p->refresh < 1 || !(p->lr > 0)) fail(DGH_ERR_ARG, "invalid training parameters");
    LearnState& L = ready_state(s);
    const auto t0 = std::chrono::steady_clock::now();
    for (int l = 1; l < s->nlevels; ++l) setup_level(s, l);
    ensure_multi_buffers(s, p->samples);
    const int nl = s->nlevels - 1;
    double first = 0.0, last = 0.0;
    int nfirst = 0;
    std::vector<double> last_round(s->nlevels, 0.0);
    for (int it = 0; it < p->steps; ++it) {
      const int l = s->nlevels - 1 - (it % nl);          // finest first, round robin
      LearnLevel& v = L.lv[l];
      if (v.S != p->samples || v.age >= p->refresh)
        gen_samples(s, l, p->samples, p->seed + 7919ull * uint64_t(l) + 104729ull * uint64_t(it));
      forward(s, l);
      const double loss = objective(s, l, v.omega.p, v.e0.p, p->samples, p->nu, true, v.gomega.p);
      if (!std::isfinite(loss)) fail(DGH_ERR_DIVERGED, "training objective is not finite at step %d", it);
      if (nfirst < nl) {
        first += loss;
        ++nfirst;
      }
      last_round[l] = loss;
      const uint32_t E = s->meshes[l]->E;
      hipLaunchKernelGGL(nodal_grad_kernel, dim3(blocks_for(size_t(E) * 8)), dim3(256), 0, s->st, E,
                         (const uint32_t*)s->meshes[l]->nl.p, (const uint32_t*)v.rowptr.p, (const float*)v.theta_n.p,
                         (const float*)v.cap.p, (const float*)v.gomega.p, v.dO.p);
      backward(s, l);
      ++L.step;
      hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(LNPARAM)), dim3(256), 0, s->st, L.theta.p, L.m1.p, L.m2.p,
                         (const float*)L.grad.p, L.wbf.p, p->lr, 0.9f, 0.999f, 1e-8f, L.step);
      HIP_CHECK(hipGetLastError());
      ++v.age;
    }
    HIP_CHECK(hipStreamSynchronize(s->st));
    for (int l = 1; l < s->nlevels; ++l) last += last_round[l];
    if (stats) {
      stats->steps = p->steps;
      stats->loss_first = nfirst ? first / nfirst : 0.0;
      stats->loss_last = nl ? last / nl : 0.0;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

int dgh_learned_apply(dgh_solver s, int enable, int samples, uint64_t seed, int* levels_on) {
  return guarded([&] {
    require_init();
    if (!s) fail(DGH_ERR_ARG, "NULL solver");
    if (enable && (s->prm.coarse_op != DGH_COARSE_JAC2 || s->dirichlet))
      fail(DGH_ERR_ARG, "learned damping needs Jacobian2 levels without boundary flags");
    refresh_operators(s);
    LearnState& L = ready_state(s);
    int mask = 0;
    if (!enable) {
      for (int l = 1; l < s->nlevels; ++l)
        if (L.on[l]) install(s, l, false);
    } else {
      check_samples(samples);
      for (int l = 1; l < s->nlevels; ++l) setup_level(s, l);
      ensure_multi_buffers(s, samples);
      const int nu = std::max(1, s->prm.nu_post);
      DevBuf<float> wj(s->meshes[s->nlevels - 1]->N);
      for (int l = 1; l < s->nlevels; ++l) {          // coarse to fine: samples see the final coarser smoothers
        if (L.on[l]) install(s, l, false);
        LearnLevel& v = L.lv[l];
        const uint32_t N = s->meshes[l]->N;
        gen_samples(s, l, samples, seed + 31ull * uint64_t(l));
        hipLaunchKernelGGL(fill_kernel<float>, dim3(blocks_for(N)), dim3(256), 0, s->st, wj.p, float(s->prm.omega), N);
        const double lj = objective(s, l, wj.p, v.e0.p, samples, nu, false, nullptr);
        forward(s, l);
        const double ll = objective(s, l, v.omega.p, v.e0.p, samples, nu, false, nullptr);
        if (ll < lj) {
          install(s, l, true);
          mask |= 1 << l;
        }
      }
    }
    drop_graphs(s);
    HIP_CHECK(hipStreamSynchronize(s->st));
    if (levels_on) *levels_on = mask;
  });
}

}  // extern "C"
