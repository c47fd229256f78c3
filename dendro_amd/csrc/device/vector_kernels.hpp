// Elementwise and reduction kernels of the smoother, the V-cycle and the outer CG.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgh {

constexpr int VBLOCK = 256;

// Richardson + Jacobi sweep after y = A x:  r = b - y;  x = x + omega (r * invd)
// (KSPSolve_Richardson with PCApply_Jacobi, bin/options:187-193).
template <class T>
__global__ void smooth_update_kernel(uint32_t n, T omega, const T* __restrict__ b,
                                     const T* __restrict__ y, const T* __restrict__ invd,
                                     T* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T r = b[i] - y[i];
  x[i] = x[i] + omega * (r * invd[i]);
}

// First sweep from a zero guess: r = b, x = omega (b * invd).
template <class T>
__global__ void smooth_first_kernel(uint32_t n, T omega, const T* __restrict__ b,
                                    const T* __restrict__ invd, T* __restrict__ x) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = T(0) + omega * (b[i] * invd[i]);
}

// r = b - y
template <class T>
__global__ void residual_kernel(uint32_t n, const T* __restrict__ b, const T* __restrict__ y,
                                T* __restrict__ r) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
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

// y = y + a x  (a read from device memory when ap != nullptr, scaled by sgn)
__global__ void axpy_kernel(uint32_t n, const double* __restrict__ ap, double a, double sgn,
                            const double* __restrict__ x, double* __restrict__ y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double al = ap ? sgn * ap[0] : a;
  y[i] = y[i] + al * x[i];
}

// p = z + b p
__global__ void aypx_kernel(uint32_t n, double b, const double* __restrict__ z,
                            double* __restrict__ p) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + b * p[i];
}

// v = v - s/n (constant-null-space projection; s = sum of v precomputed on the device)
template <class T>
__global__ void subtract_mean_kernel(uint32_t n, const double* __restrict__ sum, T* __restrict__ v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = v[i] - T(sum[0] / double(n));
}

// Wavefront (64 lanes) reduction, then one atomic per workgroup.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// out[q] += sum_i a_q[i] b_q[i] for up to 3 pairs at once (b_q == nullptr: sum of a_q;
// a_q == nullptr: pair unused).  out must be zeroed.
template <class T>
__global__ void __launch_bounds__(VBLOCK)
dot3_kernel(uint32_t n, const T* __restrict__ a0, const T* __restrict__ b0,
            const T* __restrict__ a1, const T* __restrict__ b1, const T* __restrict__ a2,
            const T* __restrict__ b2, double* __restrict__ out) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (a0) s0 += double(a0[i]) * (b0 ? double(b0[i]) : 1.0);
    if (a1) s1 += double(a1[i]) * (b1 ? double(b1[i]) : 1.0);
    if (a2) s2 += double(a2[i]) * (b2 ? double(b2[i]) : 1.0);
  }
  __shared__ double red[3][VBLOCK / 64];
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = s0;
    red[1][w] = s1;
    red[2][w] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0, t2 = 0;
    for (int q = 0; q < VBLOCK / 64; ++q) {
      t0 += red[0][q];
      t1 += red[1][q];
      t2 += red[2][q];
    }
    if (a0) atomicAdd(&out[0], t0);
    if (a1) atomicAdd(&out[1], t1);
    if (a2) atomicAdd(&out[2], t2);
  }
}

// x = Ainv b for a dense symmetric n x n inverse (row-major), one wavefront per row.
template <class T>
__global__ void __launch_bounds__(256)
dense_apply_kernel(int n, const T* __restrict__ Ainv, const T* __restrict__ b, T* __restrict__ x) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const T* a = Ainv + size_t(row) * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += double(a[j]) * double(b[j]);
  s = wave_sum(s);
  if (lane == 0) x[row] = T(s);
}

// copy the lower triangle (column-major lower == row-major upper) to a full symmetric matrix
template <class T>
__global__ void symmetrize_kernel(int n, T* __restrict__ A) {
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || i >= n || j <= i) return;
  // column-major lower: element (r=j, c=i) stored at i*n + j, row-major view (i, j)
  A[size_t(j) * n + i] = A[size_t(i) * n + j];
}

}  // namespace dgh
