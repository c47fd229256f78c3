// Elementwise and reduction kernels of the smoother, the V-cycle and the outer CG.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace dgh {

constexpr int VBLOCK = 256;

// First sweep from a zero guess: r = b, x = omega (b * invd).
template <class T>
__global__ void smooth_first_kernel(uint32_t n, T omega, const T* __restrict__ b,
                                    const T* __restrict__ invd, T* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = T(0) + omega * (b[i] * invd[i]);
}

// invd = 1 / d
template <class T>
__global__ void reciprocal_kernel(uint32_t n, const T* __restrict__ d, T* __restrict__ invd) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) invd[i] = T(1) / d[i];
}

template <class To, class Ti>
__global__ void convert_kernel(uint32_t n, const Ti* __restrict__ a, To* __restrict__ b) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = To(a[i]);
}

// Wavefront (64 lanes) reduction.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Deterministic two-stage reductions (no atomics, no zeroing, safe in graphs).
// Stage 1: block b writes partial q of sum_i a_q[i] * b_q[i] (b_q == nullptr: sum of a_q;
// a_q == nullptr: unused) to part[q * RED_BLOCKS + b].  Stage 2 sums the partials of each
// quantity in a fixed order into out[q].
constexpr int RED_BLOCKS = 1024;
constexpr int MAXQ = 4;

template <class T>
struct DotArgs {
  const T* a[MAXQ];
  const T* b[MAXQ];
};

template <class T>
__global__ void __launch_bounds__(VBLOCK)
reduce1_kernel(uint32_t n, DotArgs<T> d, double* __restrict__ part) {
  double s[MAXQ] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t i = blockIdx.x * VBLOCK + threadIdx.x; i < n; i += RED_BLOCKS * VBLOCK) {
#pragma unroll
    for (int q = 0; q < MAXQ; ++q)
      if (d.a[q]) s[q] += double(d.a[q][i]) * (d.b[q] ? double(d.b[q][i]) : 1.0);
  }
  __shared__ double red[MAXQ][VBLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const double v = wave_sum(s[q]);
    if (lane == 0) red[q][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < MAXQ) {
    double t = 0.0;
    for (int k = 0; k < VBLOCK / 64; ++k) t += red[threadIdx.x][k];
    part[threadIdx.x * RED_BLOCKS + blockIdx.x] = t;
  }
}

constexpr int RED2_THREADS = 256;
static __global__ void __launch_bounds__(RED2_THREADS)
reduce2_kernel(int nq, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[RED2_THREADS / 64];
  for (int q = 0; q < nq; ++q) {
    double v = 0.0;
    for (int k = threadIdx.x; k < RED_BLOCKS; k += RED2_THREADS) v += part[q * RED_BLOCKS + k];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int k = 0; k < RED2_THREADS / 64; ++k) t += red[k];
      out[q] = t;
    }
    __syncthreads();
  }
}

// out[0] = part[0] + ... + part[n-1] in the order cg_alpha_kernel sums the operator's tile
// partials (one workgroup of RED2_THREADS): the same value CG forms for p.w.
static __global__ void __launch_bounds__(RED2_THREADS)
sum_partials_kernel(int n, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[RED2_THREADS / 64];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += RED2_THREADS) v += part[k];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < RED2_THREADS / 64; ++k) t += red[k];
    out[0] = t;
  }
}

// r = b - Ax, and the sweep x_out = x + omega ((b - Ax) invd) written as the operator kernel's
// store phase writes it, for operators applied by other kernels (Jacobian3 levels)
template <class T>
__global__ void resid_vec_kernel(uint32_t n, const T* __restrict__ b, const T* ax, T* r) {   // r may be ax
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - ax[i];
}
template <class T>
__global__ void smooth_vec_kernel(uint32_t n, const T* __restrict__ x, const T* __restrict__ b,
                                  const T* __restrict__ ax, const T* __restrict__ invd, T omega, T* __restrict__ xo) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xo[i] = x[i] + omega * ((b[i] - ax[i]) * invd[i]);
}

// v = v - sum[0] / n  (constant-null-space projection)
template <class T>
__global__ void subtract_mean_kernel(uint32_t n, const double* __restrict__ sum, T* __restrict__ v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = v[i] - T(sum[0] / double(n));
}

// ---------------------------------------------------------------- coarsest-level CG
// Jacobi-preconditioned CG for coarsest levels too large for the dense inverse, FP64, with
// its scalars on the device: the iterations run in chunks without host synchronisation;
// once the stopping test holds (DONE) the step lengths are zero, so further iterations of
// a chunk leave x unchanged.
enum CoarseCgScalar { CC_RZ, CC_PW, CC_RZN, CC_RR, CC_RR0, CC_ALPHA, CC_BETA, CC_DONE, CC_ITS, CC_TOL2, CC_N };

// r = b (in FP64)
template <class T>
__global__ void ccg_load_kernel(uint32_t n, const T* __restrict__ b, double* __restrict__ r) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = double(b[i]);
}

// x = 0, z = invd r, p = z
static __global__ void ccg_start_kernel(uint32_t n, const double* __restrict__ invd, const double* __restrict__ r,
                                 double* __restrict__ x, double* __restrict__ z, double* __restrict__ p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double zi = invd[i] * r[i];
  x[i] = 0.0;
  z[i] = zi;
  p[i] = zi;
}

// after the first r.z and r.r: RZ, RR0, stopping tolerance, DONE for a zero right-hand side
static __global__ void ccg_init_kernel(double* S, double tol2) {
  S[CC_RZ] = S[CC_RZN];
  S[CC_RR0] = S[CC_RR];
  S[CC_TOL2] = tol2;
  S[CC_ITS] = 0.0;
  S[CC_DONE] = S[CC_RR] > 0.0 ? 0.0 : 1.0;
}

static __global__ void ccg_alpha_kernel(double* S) {
  if (S[CC_DONE] == 0.0 && !(S[CC_PW] > 0.0)) S[CC_DONE] = 1.0;   // p = 0, or not positive definite
  S[CC_ALPHA] = S[CC_DONE] == 0.0 ? S[CC_RZ] / S[CC_PW] : 0.0;
}

// x += alpha p, r -= alpha w, z = invd r
static __global__ void ccg_update_kernel(uint32_t n, const double* __restrict__ S, const double* __restrict__ p,
                                  const double* __restrict__ w, const double* __restrict__ invd,
                                  double* __restrict__ x, double* __restrict__ r, double* __restrict__ z) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = S[CC_ALPHA];
  const double ri = r[i] - a * w[i];
  x[i] = x[i] + a * p[i];
  r[i] = ri;
  z[i] = invd[i] * ri;
}

// beta = (r.z)_new / (r.z), iteration count, stopping test ||r||^2 <= tol^2 ||b||^2
static __global__ void ccg_beta_kernel(double* S) {
  if (S[CC_DONE] != 0.0) {
    S[CC_BETA] = 0.0;
    return;
  }
  S[CC_BETA] = S[CC_RZ] > 0.0 ? S[CC_RZN] / S[CC_RZ] : 0.0;
  S[CC_RZ] = S[CC_RZN];
  S[CC_ITS] += 1.0;
  if (!(S[CC_RR] > S[CC_TOL2] * S[CC_RR0])) S[CC_DONE] = 1.0;
}

// p = z + beta p
static __global__ void ccg_p_kernel(uint32_t n, const double* __restrict__ S, const double* __restrict__ z,
                             double* __restrict__ p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + S[CC_BETA] * p[i];
}

// ---------------------------------------------------------------- CG with device scalars
// S: RZ = z.r, PW = p.w, DP = stopping-test norm, RATIO = p-update factor, ZZ, ZR, ZW, RR,
// BB = b.b, ALPHA.
enum CgScalar { S_RZ = 0, S_PW, S_DP, S_RATIO, S_ZZ, S_ZR, S_ZW, S_RR, S_BB, S_ALPHA, S_PROJ, S_MU, S_NSCAL = 16 };

// p = z + ratio p  (VecAYPX)
static __global__ void cg_p_kernel(uint32_t n, const double* __restrict__ S, const double* __restrict__ z,
                            double* __restrict__ p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + S[S_RATIO] * p[i];
}

// a = rz / pw;  x = x + a p;  r = r + (-a) w  (VecAXPY twice)
static __global__ void cg_xr_kernel(uint32_t n, double* __restrict__ S, const double* __restrict__ p,
                             const double* __restrict__ w, double* __restrict__ x, double* __restrict__ r) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const double a = S[S_RZ] / S[S_PW];
  if (i == 0) S[S_ALPHA] = a;
  if (i < n) {
    x[i] = x[i] + a * p[i];
    r[i] = r[i] + (-a) * w[i];
  }
}

// ---------------------------------------------------------------- fused CG iteration
// The preconditioned residual z stays unprojected (FP32 in the mixed solver); its constant
// null-space component mu = mean(z) is applied where z is used: in the direction update
// (p' = (z - mu) + beta p, inside the operator kernel) and in the dot products, which come
// from one pass over z, r, w as
//   z'.z' = z.z - n mu^2,  z'.r = z.r - mu sum(r),  z'.w = z.w - mu sum(w).
constexpr int NDOT = 7;   // sum z, z.z, z.r, z.w, r.r, sum r, sum w
enum CgDot { D_SZ = 0, D_ZZ, D_ZR, D_ZW, D_RR, D_SR, D_SW };
constexpr int DOT_BLOCKS = 1024;  // workgroups of the dot pass (one partial each)
constexpr int DOT_UNROLL = 2;     // elements per thread and step of the dot pass
static_assert(NDOT * DOT_BLOCKS <= (MAXQ > NDOT ? MAXQ : NDOT) * RED_BLOCKS, "partials fit the buffer");

// Stage 1 of the dots: block partials part[q * DOT_BLOCKS + block] (w == nullptr: z.w and
// sum w are zero).
template <class TZ>
__global__ void __launch_bounds__(VBLOCK)
cg_dots_kernel(uint32_t n, const TZ* __restrict__ z, const double* __restrict__ r, const double* __restrict__ w,
               double* __restrict__ part) {
  double s[NDOT] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // DOT_UNROLL strided elements per step, loaded together
  constexpr uint32_t STRIDE = DOT_BLOCKS * VBLOCK;
  for (uint32_t i0 = blockIdx.x * VBLOCK + threadIdx.x; i0 < n; i0 += DOT_UNROLL * STRIDE) {
    double zi[DOT_UNROLL], ri[DOT_UNROLL], wi[DOT_UNROLL];
#pragma unroll
    for (int u = 0; u < DOT_UNROLL; ++u) {
      const uint32_t i = i0 + u * STRIDE;
      const bool in = i < n;
      zi[u] = in ? double(z[i]) : 0.0;
      ri[u] = in ? r[i] : 0.0;
      wi[u] = (in && w) ? w[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < DOT_UNROLL; ++u) {
      s[D_SZ] += zi[u];
      s[D_ZZ] += zi[u] * zi[u];
      s[D_ZR] += zi[u] * ri[u];
      s[D_ZW] += zi[u] * wi[u];
      s[D_RR] += ri[u] * ri[u];
      s[D_SR] += ri[u];
      s[D_SW] += wi[u];
    }
  }
  __shared__ double red[NDOT][VBLOCK / 64];
#pragma unroll
  for (int q = 0; q < NDOT; ++q) {
    const double v = wave_sum(s[q]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < NDOT) {
    double t = 0.0;
    for (int k = 0; k < VBLOCK / 64; ++k) t += red[threadIdx.x][k];
    part[threadIdx.x * DOT_BLOCKS + blockIdx.x] = t;
  }
}

// Sums the NDOT x DOT_BLOCKS partials in a fixed order (one RED2_THREADS-thread workgroup:
// DOT_BLOCKS / RED2_THREADS partials of every quantity per thread, all loads in flight).
static_assert(DOT_BLOCKS % RED2_THREADS == 0, "whole partials per thread");
__device__ __forceinline__ void cg_sum_dots(const double* __restrict__ part, double* out) {
  constexpr int PPT = DOT_BLOCKS / RED2_THREADS;
  __shared__ double red[NDOT][RED2_THREADS / 64];
  double v[NDOT];
  double pv[NDOT][PPT];
#pragma unroll
  for (int q = 0; q < NDOT; ++q)
#pragma unroll
    for (int k = 0; k < PPT; ++k) pv[q][k] = part[q * DOT_BLOCKS + k * RED2_THREADS + threadIdx.x];
#pragma unroll
  for (int q = 0; q < NDOT; ++q) {
    v[q] = pv[q][0];
#pragma unroll
    for (int k = 1; k < PPT; ++k) v[q] += pv[q][k];
  }
#pragma unroll
  for (int q = 0; q < NDOT; ++q) {
    v[q] = wave_sum(v[q]);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v[q];
  }
  __syncthreads();
  for (int q = 0; q < NDOT; ++q) {
    double t = 0.0;
    for (int k = 0; k < RED2_THREADS / 64; ++k) t += red[q][k];
    out[q] = t;
  }
}

// Projected dots into S (mu = 0 without a null space).
__device__ __forceinline__ void cg_projected(const double* d, uint32_t n, int null_space, double* __restrict__ S) {
  const double mu = null_space ? d[D_SZ] / double(n) : 0.0;
  S[S_MU] = mu;
  S[S_ZZ] = d[D_ZZ] - double(n) * mu * mu;
  S[S_ZR] = d[D_ZR] - mu * d[D_SR];
  S[S_ZW] = d[D_ZW] - mu * d[D_SW];
  S[S_RR] = d[D_RR];
}

// after z = B r: projected dots, dp, the next p-update factor, rz (grid: one workgroup of
// RED2_THREADS)
static __global__ void __launch_bounds__(RED2_THREADS)
cg_scalar_fused_kernel(const double* __restrict__ part, uint32_t n, int null_space, double* __restrict__ S,
                       int precond_norm, int flexible) {
  double d[NDOT];
  cg_sum_dots(part, d);
  if (threadIdx.x != 0) return;
  cg_projected(d, n, null_space, S);
  const double dp = sqrt(precond_norm ? S[S_ZZ] : S[S_RR]);
  const double ratio = flexible ? (-S[S_ALPHA] * S[S_ZW]) / S[S_RZ] : S[S_ZR] / S[S_RZ];
  S[S_DP] = dp;
  S[S_RATIO] = ratio;
  S[S_RZ] = S[S_ZR];
}

// before the first iteration (r = b projected, p = 0): rz, dp0, bb, ratio = 0
static __global__ void __launch_bounds__(RED2_THREADS)
cg_init_fused_kernel(const double* __restrict__ part, uint32_t n, int null_space, double* __restrict__ S,
                     int precond_norm) {
  double d[NDOT];
  cg_sum_dots(part, d);
  if (threadIdx.x != 0) return;
  cg_projected(d, n, null_space, S);
  S[S_BB] = d[D_RR];
  S[S_RZ] = S[S_ZR];
  S[S_DP] = sqrt(precond_norm ? S[S_ZZ] : S[S_BB]);
  S[S_RATIO] = 0.0;
}

// p = (z - mu) + beta p  (the projected z with VecAYPX)
template <class TZ>
__global__ void cg_p_fused_kernel(uint32_t n, const double* __restrict__ S, const TZ* __restrict__ z,
                                  double* __restrict__ p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (double(z[i]) - S[S_MU]) + S[S_RATIO] * p[i];
}

// p.w = the sum of the operator's tile partials (the order of sum_partials_kernel), and
// alpha = rz / p.w, into S (grid: one workgroup of RED2_THREADS)
static __global__ void __launch_bounds__(RED2_THREADS)
cg_alpha_kernel(int ntiles, const double* __restrict__ part, double* __restrict__ S) {
  __shared__ double red[RED2_THREADS / 64];
  double v = 0.0;
  for (int k = threadIdx.x; k < ntiles; k += RED2_THREADS) v += part[k];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double pw = 0.0;
    for (int k = 0; k < RED2_THREADS / 64; ++k) pw += red[k];
    S[S_PW] = pw;
    S[S_ALPHA] = S[S_RZ] / pw;
  }
}

// x = x + alpha p;  r = r + (-alpha) w;  and the V-cycle's first smoothing sweep from zero,
// x1 = omega (r invd), in the V-cycle's precision TV (mixed: also the FP32 copy of r, r32)
template <class TV>
__global__ void __launch_bounds__(VBLOCK)
cg_xr_fused_kernel(uint32_t n, const double* __restrict__ S, const double* __restrict__ p,
                   const double* __restrict__ w, double* __restrict__ x, double* __restrict__ r,
                   float* __restrict__ r32, TV omega, const TV* __restrict__ invd, TV* __restrict__ x1) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = S[S_ALPHA];
  x[i] = x[i] + a * p[i];
  const double ri = r[i] + (-a) * w[i];
  r[i] = ri;
  if (r32) r32[i] = float(ri);
  x1[i] = TV(0) + omega * (TV(ri) * invd[i]);
}

// after z = B r and the dots: dp, the next p-update factor, rz
static __global__ void cg_scalar_kernel(double* __restrict__ S, int precond_norm, int flexible) {
  if (threadIdx.x != 0) return;
  const double dp = sqrt(precond_norm ? S[S_ZZ] : S[S_RR]);
  const double ratio = flexible ? (-S[S_ALPHA] * S[S_ZW]) / S[S_RZ] : S[S_ZR] / S[S_RZ];
  S[S_DP] = dp;
  S[S_RATIO] = ratio;
  S[S_RZ] = S[S_ZR];
}

// before the first iteration: rz, dp0, ratio = 0 (p starts at zero)
static __global__ void cg_init_kernel(double* __restrict__ S, int precond_norm) {
  if (threadIdx.x != 0) return;
  S[S_RZ] = S[S_ZR];
  S[S_DP] = sqrt(precond_norm ? S[S_ZZ] : S[S_BB]);
  S[S_RATIO] = 0.0;
}

template <class T>
__global__ void copy_kernel(uint32_t n, const T* __restrict__ a, T* __restrict__ b) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Row pitch of the dense coarse inverse: a multiple of DENSE_PAD elements (16-byte aligned
// rows for the batched coarse solve's vector loads).
constexpr int DENSE_PAD = 8;

// Symmetric coarse inverse in packed lower SB x SB blocks: block (I, J), J <= I, stored
// row-major at sym_block_off(I, J), zero beyond n (diagonal blocks hold both triangles).
// y = A b reads each stored block once: block (I, J) contributes A_IJ b_J to the rows of
// block row I and, off the diagonal, A_IJ^T b_I to the rows of block row J.  The
// contributions go to partial sums Q[K][i] (K: the block index of the other factor), each
// (K, i) written exactly once, and y_i = sum_K Q[K][i] in a fixed order.
constexpr int SB = 64;
__host__ __device__ inline size_t sym_block_off(int I, int J) { return (size_t(I) * (I + 1) / 2 + J) * SB * SB; }

// (I, J) of packed block number blk = I (I + 1) / 2 + J
__device__ __forceinline__ void sym_block_ij(int blk, int& I, int& J) {
  I = int((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= blk) ++I;
  while (I * (I + 1) / 2 > blk) --I;
  J = blk - I * (I + 1) / 2;
}

// P <- packed lower blocks of the full symmetric A (row pitch ld); grid: one workgroup per
// packed block
template <class T>
__global__ void __launch_bounds__(256) sym_pack_kernel(int n, int ld, const T* __restrict__ A, T* __restrict__ P) {
  int I, J;
  sym_block_ij(blockIdx.x, I, J);
  T* dst = P + sym_block_off(I, J);
  for (int idx = threadIdx.x; idx < SB * SB; idx += 256) {
    const int r = I * SB + idx / SB, c = J * SB + idx % SB;
    dst[idx] = (r < n && c < n) ? A[size_t(r) * ld + c] : T(0);
  }
}

// Q[K][npad] <- block contributions; grid: one 128-thread workgroup per packed block.
// Wavefront 0 forms the row sums A_IJ b_J (lane = row), wavefront 1 the column sums
// A_IJ^T b_I (lane = column), both from the block staged in LDS (rows padded to SB + 1).
template <class T>
__global__ void __launch_bounds__(128)
sym_block_kernel(int n, int npad, const T* __restrict__ P, const T* __restrict__ b, double* __restrict__ Q) {
  __shared__ T s[SB][SB + 1];
  __shared__ T bi[SB], bj[SB];
  int I, J;
  sym_block_ij(blockIdx.x, I, J);
  const T* src = P + sym_block_off(I, J);
  // 16-byte loads, all of a thread's in flight together (a block is 16 KB in FP32)
  using V4 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  constexpr int PER = int(sizeof(V4) / sizeof(T));
  constexpr int NV = SB * SB / PER / 128;
  V4 v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = reinterpret_cast<const V4*>(src)[threadIdx.x + q * 128];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const int idx = (threadIdx.x + q * 128) * PER;
    const T* e = reinterpret_cast<const T*>(&v[q]);
#pragma unroll
    for (int u = 0; u < PER; ++u) s[(idx + u) / SB][(idx + u) % SB] = e[u];
  }
  if (threadIdx.x < SB) {
    const int r = I * SB + threadIdx.x;
    bi[threadIdx.x] = r < n ? b[r] : T(0);
  } else {
    const int c = J * SB + threadIdx.x - SB;
    bj[threadIdx.x - SB] = c < n ? b[c] : T(0);
  }
  __syncthreads();
  if (threadIdx.x < SB) {
    const int r = threadIdx.x;
    double u = 0.0;
#pragma unroll 16
    for (int c = 0; c < SB; ++c) u += double(s[r][c]) * double(bj[c]);
    Q[size_t(J) * npad + size_t(I) * SB + r] = u;
  } else if (I != J) {
    const int c = threadIdx.x - SB;
    double v = 0.0;
#pragma unroll 16
    for (int r = 0; r < SB; ++r) v += double(s[r][c]) * double(bi[r]);
    Q[size_t(I) * npad + size_t(J) * SB + c] = v;
  }
}

// y_i = sum_{K < nb} Q[K][i] - mu c1_i: one 256-thread workgroup per 64 rows, the four
// wavefronts take every fourth K and combine in LDS (fixed order).  With a constant null
// space (c1 != nullptr) the right-hand side is projected inside the solve:
// Ainv (b - mu 1) = Ainv b - mu Ainv 1 with mu = mean(b) and c1 = Ainv 1 (setup); every
// workgroup forms mu from b in the same fixed order.
template <class T>
__global__ void __launch_bounds__(256)
sym_sum_kernel(int n, int nb, int npad, const double* __restrict__ Q, const T* __restrict__ b,
               const double* __restrict__ c1, T* __restrict__ y) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // loads are issued SU at a time and added in index order (the same sums as one at a time)
  constexpr int SU = 8;
  double mu = 0.0;
  if (c1) {
    double sb = 0.0;
    for (int j0 = threadIdx.x; j0 < n; j0 += 256 * SU) {
      double v[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) v[u] = j0 + 256 * u < n ? double(b[j0 + 256 * u]) : 0.0;
#pragma unroll
      for (int u = 0; u < SU; ++u) sb += v[u];
    }
    sb = wave_sum(sb);
    if (lane == 0) red[wv][0] = sb;
    __syncthreads();
    mu = ((red[0][0] + red[1][0]) + (red[2][0] + red[3][0])) / double(n);
    __syncthreads();
  }
  const int i = blockIdx.x * 64 + lane;
  double acc = 0.0;
  if (i < npad) {
    for (int K0 = wv; K0 < nb; K0 += 4 * SU) {
      double v[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) v[u] = K0 + 4 * u < nb ? Q[size_t(K0 + 4 * u) * npad + i] : 0.0;
#pragma unroll
      for (int u = 0; u < SU; ++u) acc += v[u];
    }
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && i < n) {
    const double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    y[i] = T(c1 ? v - mu * c1[i] : v);
  }
}

// c1 = A 1 (row sums of the full symmetric A), one wavefront per row
template <class T>
__global__ void __launch_bounds__(256) row_sum_kernel(int n, int ld, const T* __restrict__ A, double* __restrict__ c1) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  double t = 0.0;
  for (int j = lane; j < n; j += 64) t += double(A[size_t(row) * ld + j]);
  t = wave_sum(t);
  if (lane == 0) c1[row] = t;
}

// copy the lower triangle (column-major lower == row-major upper) to a full symmetric matrix
// of pitch ld
template <class T>
__global__ void symmetrize_kernel(int n, int ld, T* __restrict__ A) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || i >= n || j <= i) return;
  // column-major lower: element (r=j, c=i) stored at i*ld + j, row-major view (i, j)
  A[size_t(j) * ld + i] = A[size_t(i) * ld + j];
}

}  // namespace dgh
