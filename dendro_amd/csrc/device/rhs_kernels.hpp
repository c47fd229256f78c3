// Kernels that evaluate functions at element points: quadrature right-hand sides, nodal values
// and L2 errors (the reference's ComputeRHS2 / ComputeSol / ComputeError5 loops,
// examples/src/backend/omgRhs.C).  One thread per element; elements are located by their DA
// anchor and level (int4 {x, y, z, lev}); the unit cube is DA coordinates / 2^(maxD-1).
#pragma once

#include <hip/hip_runtime.h>

#include "dendro_hip.h"
#include "vector_kernels.hpp"

namespace dgh {

constexpr int NQ = 27;   // 3x3x3 Gauss nodes, index 9 m + 3 n + p (m along x)

struct FnDev {
  int kind, nfreq;
  double a0, a1;
  double c[8];
};

__device__ __forceinline__ double eval_fn(const FnDev& f, double x, double y, double z) {
  if (f.kind == DGH_FN_GAUSSIANS) {
    double r = 0.0;
    if (x >= 0.0 && y >= 0.0 && z >= 0.0 && x < 0.2 && y < 0.2 && z < 0.2) {
      const double dx = x - 0.1, dy = y - 0.1, dz = z - 0.1;
      r += exp(-((dx * dx + dy * dy + dz * dz) / 0.005));
    }
    if (x >= 0.2 && y >= 0.2 && z >= 0.2 && x <= 0.4 && y <= 0.4 && z <= 0.4) {
      const double dx = x - 0.3, dy = y - 0.3, dz = z - 0.3;
      r += exp(-((dx * dx + dy * dy + dz * dz) / 0.005));
    }
    if (x >= 0.5 && y >= 0.5 && z >= 0.5 && x <= 0.7 && y <= 0.7 && z <= 0.7) {
      const double dx = x - 0.6, dy = y - 0.6, dz = z - 0.6;
      r += exp(-((dx * dx + dy * dy + dz * dz) / 0.005));
    }
    return r;
  }
  if (f.kind == DGH_FN_COS_SUM) {
    const double pi = 3.14159265358979323846;
    double r = 0.0;
    for (int k = 1; k <= f.nfreq; ++k) {
      const double kk = double(k);
      r += (f.a0 + 3.0 * f.a1 * kk * kk * pi * pi) * cos(kk * pi * x) * cos(kk * pi * y) * cos(kk * pi * z);
    }
    return r;
  }
  return f.c[0] + f.c[1] * x + f.c[2] * y + f.c[3] * z + f.c[4] * x * y + f.c[5] * y * z + f.c[6] * z * x +
         f.c[7] * x * y * z;
}

// Gauss node t (0..2) of the 3-point rule on [-1, 1]: {0, +sqrt(3/5), -sqrt(3/5)}, weights
// {8/9, 5/9, 5/9}
__device__ __forceinline__ double gauss_pt(int t) {
  return t == 0 ? 0.0 : (t == 1 ? 0.77459666924148337704 : -0.77459666924148337704);
}
__device__ __forceinline__ double gauss_wt(int t) { return t == 0 ? 8.0 / 9.0 : 5.0 / 9.0; }

struct ElemGeo {
  double x, y, z, h;
};
__device__ __forceinline__ ElemGeo elem_geo(int4 a, uint32_t maxD) {
  const double unit = 1.0 / double(1u << (maxD - 1));
  return ElemGeo{double(a.x) * unit, double(a.y) * unit, double(a.z) * unit,
                 double(1u << (maxD - uint32_t(a.w))) * unit};
}

// b[idx[j]] += h^3/8 sum_q w_q f(x_q) S[s][j][q]  (b zeroed by the caller)
__global__ void __launch_bounds__(VBLOCK)
rhs_quad_kernel(uint32_t E, uint32_t maxD, const int4* __restrict__ geo, const uint32_t* __restrict__ nl,
                const uint8_t* __restrict__ cs, const double* __restrict__ SF, FnDev f, double* __restrict__ b) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const ElemGeo G = elem_geo(geo[e], maxD);
  double fw[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int m = q / 9, n = (q / 3) % 3, p = q % 3;
    const double X = G.h * (1.0 + gauss_pt(m)) * 0.5 + G.x;
    const double Y = G.h * (1.0 + gauss_pt(n)) * 0.5 + G.y;
    const double Z = G.h * (1.0 + gauss_pt(p)) * 0.5 + G.z;
    fw[q] = gauss_wt(m) * gauss_wt(n) * gauss_wt(p) * eval_fn(f, X, Y, Z);
  }
  const double fac = G.h * G.h * G.h / 8.0;
  const double* S = SF + size_t(cs[e]) * 8 * NQ;
  for (int j = 0; j < 8; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc += fw[q] * S[j * NQ + q];
    atomicAdd(&b[nl[size_t(e) * 8 + j]], fac * acc);
  }
}

// u[node] = f(node) over the non-hanging vertices of every element (equal values when written
// twice)
__global__ void __launch_bounds__(VBLOCK)
nodal_values_kernel(uint32_t E, uint32_t maxD, const int4* __restrict__ geo, const uint32_t* __restrict__ nl,
                    const uint8_t* __restrict__ hn, FnDev f, double* __restrict__ u) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const ElemGeo G = elem_geo(geo[e], maxD);
  const uint8_t mask = hn[e];
  for (int k = 0; k < 8; ++k) {
    if (mask & (1u << k)) continue;
    u[nl[size_t(e) * 8 + k]] =
        eval_fn(f, G.x + double(k & 1) * G.h, G.y + double((k >> 1) & 1) * G.h, G.z + double((k >> 2) & 1) * G.h);
  }
}

// part[block] = sum over the block's elements of h^3/8 sum_q w_q (f(x_q) - u_h(x_q))^2, in a
// fixed order (grid-stride over elements, wavefront sums, then the block)
__global__ void __launch_bounds__(VBLOCK)
l2_error_kernel(uint32_t E, uint32_t maxD, const int4* __restrict__ geo, const uint32_t* __restrict__ nl,
                const uint8_t* __restrict__ cs, const double* __restrict__ SF, FnDev f, const double* __restrict__ u,
                double* __restrict__ part) {
  double tot = 0.0;
  for (uint32_t e = blockIdx.x * VBLOCK + threadIdx.x; e < E; e += gridDim.x * VBLOCK) {
    const ElemGeo G = elem_geo(geo[e], maxD);
    double ue[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) ue[j] = u[nl[size_t(e) * 8 + j]];
    const double* S = SF + size_t(cs[e]) * 8 * NQ;
    double acc = 0.0;
    for (int q = 0; q < NQ; ++q) {
      const int m = q / 9, n = (q / 3) % 3, p = q % 3;
      const double X = G.h * (1.0 + gauss_pt(m)) * 0.5 + G.x;
      const double Y = G.h * (1.0 + gauss_pt(n)) * 0.5 + G.y;
      const double Z = G.h * (1.0 + gauss_pt(p)) * 0.5 + G.z;
      double uh = 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j) uh += ue[j] * S[j * NQ + q];
      const double d = eval_fn(f, X, Y, Z) - uh;
      acc += gauss_wt(m) * gauss_wt(n) * gauss_wt(p) * d * d;
    }
    tot += acc * (G.h * G.h * G.h / 8.0);
  }
  __shared__ double red[VBLOCK / 64];
  tot = wave_sum(tot);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < VBLOCK / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

}  // namespace dgh
