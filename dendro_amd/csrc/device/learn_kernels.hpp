// Kernels of the learned Jacobi damping: a per-element MLP (35 features -> 64 -> 64 -> 8
// damping logits, ReLU) with BF16 weights and FP32 accumulation, its backward pass with
// deterministic weight-gradient reduction, Adam, the nodal averaging of the logits into
// damping weights, and the batched smoothing objective.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_multi.hpp"

namespace dgh {

constexpr int LNF = 35, LNH = 64, LNO = 8;
constexpr int L_W1 = 0, L_B1 = L_W1 + LNH * LNF, L_W2 = L_B1 + LNH, L_B2 = L_W2 + LNH * LNH,
              L_W3 = L_B2 + LNH, L_B3 = L_W3 + LNO * LNH, LNPARAM = L_B3 + LNO;
constexpr int LTE = 64;          // elements per tile
constexpr int LBLOCK = 256;
constexpr int LXS = LNF + 2;     // odd LDS row strides: column walks hit distinct banks
constexpr int LHS = LNH + 1;

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// round to nearest even
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// element arrays of one level
struct LearnElems {
  const uint32_t* nl;    // [E][8]
  const uint8_t* cs;     // [E] 18 child + type
  const uint8_t* lev;    // [E]
  const float* lke;      // [E] log10 kappa_e
  const float* lkbar;    // [N] log10 of the nodal mean coefficient
  uint32_t E, maxD;
};

// Forward (BWD = false: O = MLP(X)) or forward + backward (BWD = true: per-workgroup
// partial gradients of sum_e dO[e] . O[e] with respect to all parameters into
// gpart[blockIdx.x][LNPARAM]).  Workgroups walk tiles of LTE elements.
template <bool BWD>
__global__ void __launch_bounds__(LBLOCK)
mlp_kernel(LearnElems le, const uint16_t* __restrict__ wbf, const float* __restrict__ theta,
           float* __restrict__ O, const float* __restrict__ dO, float* __restrict__ gpart) {
  __shared__ float sW1[LNH * LNF], sW2[LNH * LNH], sW3[LNO * LNH], sb1[LNH], sb2[LNH], sb3[LNO];
  __shared__ float sX[LTE * LXS], sH1[LTE * LHS], sH2[LTE * LHS];
  __shared__ float sD1[BWD ? LTE * LHS : 1], sD2[BWD ? LTE * LHS : 1], sdO[BWD ? LTE * LNO : 1];
  const int t = threadIdx.x;
  for (int i = t; i < LNH * LNF; i += LBLOCK) sW1[i] = bf16_to_f32(wbf[L_W1 + i]);
  for (int i = t; i < LNH * LNH; i += LBLOCK) sW2[i] = bf16_to_f32(wbf[L_W2 + i]);
  for (int i = t; i < LNO * LNH; i += LBLOCK) sW3[i] = bf16_to_f32(wbf[L_W3 + i]);
  if (t < LNH) {
    sb1[t] = theta[L_B1 + t];
    sb2[t] = theta[L_B2 + t];
  }
  if (t < LNO) sb3[t] = theta[L_B3 + t];
  // per-thread gradient accumulators (BWD): W2 4x4 block, W1 row segment, W3 pair, one bias
  float gw2[4][4] = {}, gw1[9] = {}, gw3[2] = {}, gb = 0.f;
  const int e_lo = t & 63, grp = t >> 6;
  const int jr = (t >> 4) * 4, ic = (t & 15) * 4;            // W2 block
  const int j1 = t >> 2, c0 = (t & 3) * 9, nc1 = (t & 3) == 3 ? 8 : 9;   // W1 row segment
  const int m3 = t >> 5, j3 = (t & 31) * 2;                  // W3 pair
  const uint32_t ntiles = (le.E + LTE - 1) / LTE;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t e0 = tile * LTE;
    __syncthreads();                       // previous tile (and the weight staging) done
    if (t < LTE) {
      float* x = sX + t * LXS;
      for (int i = 0; i < LNF; ++i) x[i] = 0.f;
      const uint32_t e = e0 + t;
      if (e < le.E) {
        const uint32_t cs = le.cs[e];
        x[0] = float(double(le.lev[e] - 1) / double(le.maxD - 1));
        x[1 + cs % 18] = 1.f;
        x[19 + cs / 18] = 1.f;
        const float lk = le.lke[e];
        for (int k = 0; k < 8; ++k) x[27 + k] = le.lkbar[le.nl[size_t(e) * 8 + k]] - lk;
      }
    }
    __syncthreads();
    for (int u = 0; u < 16; ++u) {
      const int j = grp * 16 + u;
      float a = sb1[j];
      for (int i = 0; i < LNF; ++i) a += sW1[j * LNF + i] * sX[e_lo * LXS + i];
      sH1[e_lo * LHS + j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int u = 0; u < 16; ++u) {
      const int j = grp * 16 + u;
      float a = sb2[j];
      for (int i = 0; i < LNH; ++i) a += sW2[j * LNH + i] * sH1[e_lo * LHS + i];
      sH2[e_lo * LHS + j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (!BWD) {
      for (int u = 0; u < 2; ++u) {
        const int m = grp * 2 + u;
        float a = sb3[m];
        for (int j = 0; j < LNH; ++j) a += sW3[m * LNH + j] * sH2[e_lo * LHS + j];
        if (e0 + e_lo < le.E) O[size_t(e0 + e_lo) * LNO + m] = a;
      }
      continue;
    }
    for (int i = t; i < LTE * LNO; i += LBLOCK) {
      const uint32_t e = e0 + i / LNO;
      sdO[i] = e < le.E ? dO[size_t(e0) * LNO + i] : 0.f;
    }
    __syncthreads();
    for (int u = 0; u < 16; ++u) {
      const int j = grp * 16 + u;
      float a = 0.f;
      for (int m = 0; m < LNO; ++m) a += sW3[m * LNH + j] * sdO[e_lo * LNO + m];
      sD2[e_lo * LHS + j] = sH2[e_lo * LHS + j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int u = 0; u < 16; ++u) {
      const int i = grp * 16 + u;
      float a = 0.f;
      for (int j = 0; j < LNH; ++j) a += sW2[j * LNH + i] * sD2[e_lo * LHS + j];
      sD1[e_lo * LHS + i] = sH1[e_lo * LHS + i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int e = 0; e < LTE; ++e) {
      float d2[4], h1[4];
      for (int a = 0; a < 4; ++a) {
        d2[a] = sD2[e * LHS + jr + a];
        h1[a] = sH1[e * LHS + ic + a];
      }
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) gw2[a][b] += d2[a] * h1[b];
      const float d1 = sD1[e * LHS + j1];
      for (int c = 0; c < 9; ++c)
        if (c < nc1) gw1[c] += d1 * sX[e * LXS + c0 + c];
      const float go = sdO[e * LNO + m3];
      gw3[0] += go * sH2[e * LHS + j3];
      gw3[1] += go * sH2[e * LHS + j3 + 1];
      if (t < LNH) gb += sD1[e * LHS + t];
      else if (t < 2 * LNH) gb += sD2[e * LHS + t - LNH];
      else if (t < 2 * LNH + LNO) gb += sdO[e * LNO + t - 2 * LNH];
    }
  }
  if (BWD) {
    float* g = gpart + size_t(blockIdx.x) * LNPARAM;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) g[L_W2 + (jr + a) * LNH + ic + b] = gw2[a][b];
    for (int c = 0; c < 9; ++c)
      if (c < nc1) g[L_W1 + j1 * LNF + c0 + c] = gw1[c];
    g[L_W3 + m3 * LNH + j3] = gw3[0];
    g[L_W3 + m3 * LNH + j3 + 1] = gw3[1];
    if (t < LNH) g[L_B1 + t] = gb;
    else if (t < 2 * LNH) g[L_B2 + t - LNH] = gb;
    else if (t < 2 * LNH + LNO) g[L_B3 + t - 2 * LNH] = gb;
  }
}

// grad[p] = sum over workgroups (fixed order) of the partials
__global__ void __launch_bounds__(256)
grad_reduce_kernel(int nblocks, const float* __restrict__ gpart, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= LNPARAM) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += gpart[size_t(b) * LNPARAM + p];
  grad[p] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; refreshes the BF16 weights.
__global__ void __launch_bounds__(256)
adam_kernel(float* __restrict__ theta, float* __restrict__ m1, float* __restrict__ m2,
            const float* __restrict__ grad, uint16_t* __restrict__ wbf, float lr, float b1, float b2,
            float eps, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= LNPARAM) return;
  const float g = grad[p];
  const float a = b1 * m1[p] + (1.f - b1) * g;
  const float v = b2 * m2[p] + (1.f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float ah = a / (1.f - powf(b1, float(step))), vh = v / (1.f - powf(b2, float(step)));
  const float th = theta[p] - lr * ah / (sqrtf(vh) + eps);
  theta[p] = th;
  wbf[p] = f32_to_bf16(th);
}

__global__ void __launch_bounds__(256)
to_bf16_kernel(int n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) b[p] = f32_to_bf16(a[p]);
}

// theta_n = mean of the element logits at node n (CSR of (e, k) slots in (e, k) order),
// omega_n = cap_n sigmoid(theta_n), scaled inverse diagonals for the smoother.
__global__ void __launch_bounds__(256)
nodal_kernel(uint32_t N, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ slot,
             const float* __restrict__ O, const float* __restrict__ cap, const float* __restrict__ dinv,
             float* __restrict__ theta_n, float* __restrict__ omega, float* __restrict__ sinv32,
             double* __restrict__ sinv64, const double* __restrict__ dinv64) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const uint32_t p0 = rowptr[n], p1 = rowptr[n + 1];
  float s = 0.f;
  for (uint32_t p = p0; p < p1; ++p) s += O[slot[p]];
  const float th = p1 > p0 ? s / float(p1 - p0) : 0.f;
  const float w = cap[n] / (1.f + expf(-th));
  theta_n[n] = th;
  omega[n] = w;
  if (sinv32) sinv32[n] = w * dinv[n];
  if (sinv64) sinv64[n] = double(w) * dinv64[n];
}

// dO[e][k] = g_omega[n] cap[n] s (1 - s) / cnt[n], n = nl[e][k], s = sigmoid(theta_n)
__global__ void __launch_bounds__(256)
nodal_grad_kernel(uint32_t E, const uint32_t* __restrict__ nl, const uint32_t* __restrict__ rowptr,
                  const float* __restrict__ theta_n, const float* __restrict__ cap,
                  const float* __restrict__ g_omega, float* __restrict__ dO) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(E) * 8) return;
  const uint32_t n = nl[i];
  const float sg = 1.f / (1.f + expf(-theta_n[n]));
  dO[i] = g_omega[n] * cap[n] * sg * (1.f - sg) / float(rowptr[n + 1] - rowptr[n]);
}

// ---------------------------------------------------------------- smoothing objective
// one damped sweep on S columns: z = dinv (A e), e_next = e - omega z
__global__ void __launch_bounds__(256)
lsweep_kernel(uint32_t n, int S, const float* __restrict__ dinv, const float* __restrict__ omega,
              const float* __restrict__ Ae, const float* __restrict__ e, float* __restrict__ z,
              float* __restrict__ e_next) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(n) * S) return;
  const size_t nd = i / S;
  const float zz = dinv[nd] * Ae[i];
  z[i] = zz;
  e_next[i] = e[i] - omega[nd] * zz;
}

// g = 2 (A e_nu) / (|e_0|_A^2 S)
__global__ void __launch_bounds__(256)
lgrad_init_kernel(uint32_t n, int S, const float* __restrict__ Aef, const double* __restrict__ ea,
                  float* __restrict__ g) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(n) * S) return;
  g[i] = float(2.0 * double(Aef[i]) / (ea[i % S] * double(S)));
}

// g_omega[n] (+)= sum_s -g[n][s] z[n][s];  u = dinv omega g
__global__ void __launch_bounds__(256)
lgrad_step_kernel(uint32_t n, int S, int accumulate, const float* __restrict__ g, const float* __restrict__ z,
                  const float* __restrict__ dinv, const float* __restrict__ omega, float* __restrict__ g_omega,
                  float* __restrict__ u) {
  const uint32_t nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= n) return;
  float acc = 0.f;
  const float sc = dinv[nd] * omega[nd];
  for (int s = 0; s < S; ++s) {
    const size_t i = size_t(nd) * S + s;
    acc += -g[i] * z[i];
    u[i] = sc * g[i];
  }
  g_omega[nd] = accumulate ? g_omega[nd] + acc : acc;
}

// a -= b
__global__ void __launch_bounds__(256) sub_kernel(size_t n, float* __restrict__ a, const float* __restrict__ b) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] -= b[i];
}

// a = b - c
__global__ void __launch_bounds__(256)
diff_kernel(size_t n, const float* __restrict__ b, const float* __restrict__ c, float* __restrict__ a) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i] - c[i];
}

// uniform [-1, 1) from a counter-based hash (splitmix64 of seed and index)
__global__ void __launch_bounds__(256) rand_kernel(size_t n, uint64_t seed, float* __restrict__ v) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t(i) + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  v[i] = float(double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

}  // namespace dgh
