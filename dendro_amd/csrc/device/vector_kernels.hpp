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

// v = v - sum[0] / n  (constant-null-space projection)
template <class T>
__global__ void subtract_mean_kernel(uint32_t n, const double* __restrict__ sum, T* __restrict__ v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = v[i] - T(sum[0] / double(n));
}

// ---------------------------------------------------------------- CG with device scalars
// S: RZ = z.r, PW = p.w, DP = stopping-test norm, RATIO = p-update factor, ZZ, ZR, ZW, RR,
// BB = b.b, ALPHA.
enum CgScalar { S_RZ = 0, S_PW, S_DP, S_RATIO, S_ZZ, S_ZR, S_ZW, S_RR, S_BB, S_ALPHA, S_PROJ, S_NSCAL = 16 };

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

// x = Ainv b for a dense symmetric n x n inverse stored row-major with row pitch ld (a
// multiple of DENSE_PAD elements, 16-byte aligned rows), one wavefront per row.  Each lane
// streams 16-byte vectors, four in flight; sums in FP64 in a fixed order.
constexpr int DENSE_PAD = 8;
template <class T>
__global__ void __launch_bounds__(256)
dense_apply_kernel(int n, int ld, const T* __restrict__ Ainv, const T* __restrict__ b, T* __restrict__ x) {
  constexpr int VEC = 16 / sizeof(T);
  using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const V* a = reinterpret_cast<const V*>(Ainv + size_t(row) * ld);
  const V* bv = reinterpret_cast<const V*>(b);
  const int nv = n / VEC;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int v0 = lane; v0 < nv; v0 += 256) {
    V av[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int v = v0 + 64 * u;
      if (v < nv) { av[u] = a[v]; xv[u] = bv[v]; }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (v0 + 64 * u < nv) {
        const T* pa = reinterpret_cast<const T*>(&av[u]);
        const T* px = reinterpret_cast<const T*>(&xv[u]);
#pragma unroll
        for (int c = 0; c < VEC; ++c) s[u] += double(pa[c]) * double(px[c]);
      }
    }
  }
  double t = (s[0] + s[1]) + (s[2] + s[3]);
  const int j = nv * VEC + lane;    // tail
  if (j < n) t += double(Ainv[size_t(row) * ld + j]) * double(b[j]);
  t = wave_sum(t);
  if (lane == 0) x[row] = T(t);
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
