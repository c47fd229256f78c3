// Linear elasticity operator (3 dofs per node) of libdendrohip: the reference's
// ElasticityMatMult (examples/src/backend/elasticityJac.C:129-211),
//   y_k = sum_j fac (mu L[k][j] I + (mu + lambda) GD[3k..3k+2][3j..3j+2]) x_j,  fac = h/2,
// with optional Dirichlet flags (boundary rows y = x, boundary columns dropped).
//
// Owner-computes tiles as for the scalar operator (tiles.cpp plan; one workgroup per tile):
// the tile's touched nodes' three components are staged in LDS, every record (own and halo
// elements) is evaluated, the rows of the nodes the tile owns are accumulated with 64-bit
// fixed-point LDS additions, and each owned node is stored once (no global atomics, no
// zeroing of y, bit-reproducible).
// Element matrices: child c's element is child 0's with the vertex order k^c (the plan's
// vertex permutation) and the signs s_a s_b on the (a, b) blocks (s_a = -1 when bit a of c is
// set: the mirror flips axis a).  Child 0's element of hanging mask H is the regular one
// constrained by the interpolation P_H of its hanging vertices, K_H = P_H^T K_0 P_H, where
// row k of P_H is e_k for a regular vertex, (e_0 + e_k)/2 for a hanging edge midpoint
// (k = 1, 2, 4) and (e_0 + e_a + e_b + e_k)/4 for a hanging face centre (k = a | b =
// 3, 5, 6); the stencils of the 18 child-0 types are exactly these products (the shape
// functions interpolate the parent's vertices).  So a lane applies P_H to its (sign-flipped)
// inputs, the one 24x24 matrix K_0 = mu L_0 (x) I + (mu + lambda) GD_0 of the unit element
// (every lane the same: wave-uniform rows through the scalar cache, no per-type passes), and
// P_H^T to the result, and scales by its element's fac = h/2.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "vector_kernels.hpp"

namespace dgh {

constexpr int ELAS_ROWS = 24;
constexpr int ELAS_NSTRIDE = 24;   // LDS bytes per node: three doubles (sums or staged x)
constexpr int ELAS_XCHUNK = 4;     // node gathers per thread in flight together

// Threads per tile, owned-node capacity and node budget (owned capacity + touched nodes) of an
// elasticity tile: the kernel's registers (about 160 VGPRs) allow three 256-thread workgroups
// per CU, and the budget gives each the LDS of a third of a CU (24 B per node slot).  On the
// 2M-element log-normal octree: 659 us per product for a two-pass (element outputs to HBM,
// node gather) version, 387 us for per-(type, level) 24x24 class matrices, 234 us for the
// P_H^T K_0 P_H form with 512-thread tiles of 1680 node slots, 182 us with 256 threads and
// 170 us with this budget.
constexpr int ELAS_BLOCK = 256;
constexpr uint32_t ELAS_OWN_CAP = 640u;
constexpr uint32_t ELAS_BUDGET = 2200u;
constexpr uint32_t ELAS_XS_OFF = ELAS_NSTRIDE * ELAS_OWN_CAP + 64u;
__host__ constexpr size_t elas_lds_bytes(uint32_t max_touch) {
  return size_t(ELAS_XS_OFF) + size_t(ELAS_NSTRIDE) * max_touch;
}

// The unit matrix is read through the constant address space: uniform loads from it are
// scalar loads whatever the address arithmetic.
template <class U>
using const_ptr = const __attribute__((address_space(4))) U*;

__device__ __forceinline__ double flip_sign(double v, uint32_t neg) {
  return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<long long>(neg) << 63));
}

// meta: child-0 hanging mask | child << 8 | level << 11.  K0: the unit element's [24][24]
// (rows 3k + a); kzero: 0.  texp[t]: exponent bound of the tile's row sums of |fac K_H| and log2 bound of
// the terms per node (as for the scalar operator).
__global__ void __launch_bounds__(ELAS_BLOCK)
elas_tile_kernel(const TileRec* __restrict__ tiles, int ntiles, const uint16_t* __restrict__ lid,
                 const uint16_t* __restrict__ meta, const uint32_t* __restrict__ node_tab,
                 const int2* __restrict__ texp, const double* __restrict__ K0, uint32_t kzero, uint32_t maxD,
                 double hFac, const uint8_t* __restrict__ bdy, const double* __restrict__ x, double* __restrict__ y) {
  constexpr int BLOCK = ELAS_BLOCK;
  constexpr int NW = BLOCK / 64;
  constexpr uint32_t XS = ELAS_XS_OFF;
  static_assert(NW <= 16, "one max |x| word per wavefront in 64 LDS bytes");
  static_assert(XS + 16u <= 65535u, "staged x must start within the ds immediate offset range");
  const lds_ptr<uint32_t> s_xhi = lds_at<uint32_t>(ELAS_NSTRIDE * ELAS_OWN_CAP);
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  const TileRec rec = tiles[t];
  const uint32_t* nodes = node_tab + rec.node_off;
  const uint32_t n_own = rec.n_own;
  uint4 lp = make_uint4(0, 0, 0, 0);
  uint32_t mt = 0;
  if (threadIdx.x < rec.n_elem) {
    lp = reinterpret_cast<const uint4*>(lid)[rec.elem_off + threadIdx.x];
    mt = meta[rec.elem_off + threadIdx.x];
  }
  // stage the touched nodes' x (boundary nodes as 0: their columns drop out)
  uint32_t xhi = 0;
  for (uint32_t base = 0; base < rec.n_touch; base += ELAS_XCHUNK * BLOCK) {
    uint32_t nid[ELAS_XCHUNK];
#pragma unroll
    for (int q = 0; q < ELAS_XCHUNK; ++q) {
      const uint32_t i = base + threadIdx.x + q * BLOCK;
      nid[q] = (i < rec.n_touch) ? nodes[i] : 0u;
    }
    double v[ELAS_XCHUNK][3];
#pragma unroll
    for (int q = 0; q < ELAS_XCHUNK; ++q) {
      const uint32_t i = base + threadIdx.x + q * BLOCK;
      const bool in = i < rec.n_touch;
#pragma unroll
      for (int a = 0; a < 3; ++a) v[q][a] = in ? x[3 * size_t(nid[q]) + a] : 0.0;
    }
    if (bdy) {
#pragma unroll
      for (int q = 0; q < ELAS_XCHUNK; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        if (i < rec.n_touch && bdy[nid[q]]) v[q][0] = v[q][1] = v[q][2] = 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < ELAS_XCHUNK; ++q) {
      const uint32_t i = base + threadIdx.x + q * BLOCK;
      if (i < rec.n_touch) {
#pragma unroll
        for (int a = 0; a < 3; ++a) *lds_at<double>(XS + ELAS_NSTRIDE * i + 8u * a) = v[q][a];
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
        xhi = max(xhi, uint32_t(uint64_t(__double_as_longlong(v[q][a])) >> 32) & 0x7fffffffu);
    }
  }
  for (uint32_t i = threadIdx.x; i < 3 * n_own; i += BLOCK) lds_at<unsigned long long>(0)[i] = 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) xhi = max(xhi, uint32_t(__shfl_xor(int(xhi), off)));
  if ((threadIdx.x & 63) == 0) s_xhi[threadIdx.x >> 6] = xhi;
  __syncthreads();
#pragma unroll
  for (int w = 0; w < NW; ++w) xhi = max(xhi, s_xhi[w]);
  const bool finite = (xhi >> 20) < 0x7ffu;
  const int ex = max(int(xhi >> 20) - 1022, FX_XMIN);
  const int es = fx_scale<double>(texp[t], ex);
  const uint32_t own_bytes = n_own * ELAS_NSTRIDE;
  for (uint32_t base = 0;;) {
    if (base + threadIdx.x < rec.n_elem) {
      uint32_t l[8];
      l[0] = (lp.x & 0xffffu) * ELAS_NSTRIDE; l[1] = (lp.x >> 16) * ELAS_NSTRIDE;
      l[2] = (lp.y & 0xffffu) * ELAS_NSTRIDE; l[3] = (lp.y >> 16) * ELAS_NSTRIDE;
      l[4] = (lp.z & 0xffffu) * ELAS_NSTRIDE; l[5] = (lp.z >> 16) * ELAS_NSTRIDE;
      l[6] = (lp.w & 0xffffu) * ELAS_NSTRIDE; l[7] = (lp.w >> 16) * ELAS_NSTRIDE;
      const uint32_t hm = mt & 0xffu, c = (mt >> 8) & 7u, lev = mt >> 11;
      const uint32_t neg[3] = {c & 1u, (c >> 1) & 1u, (c >> 2) & 1u};
      // the element's fac = h/2 and 2^es in one scale per component (with the mirror sign)
      const double sc = ldexp(hFac * double(1u << (maxD - lev)) / 2.0, es);
      const double sa[3] = {flip_sign(sc, neg[0]), flip_sign(sc, neg[1]), flip_sign(sc, neg[2])};
      double xe[ELAS_ROWS];
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) xe[3 * k + a] = flip_sign(*lds_at<double>(XS + l[k] + 8u * a), neg[a]);
      // per hanging vertex k: c_k = the weight of x_k in x'_k, d_k = the weight of each other
      // value in it (edge k = 1, 2, 4: (x_0 + x_k)/2; face k = 3, 5, 6: (x_0 + x_a + x_b + x_k)/4)
      double ck[7], dk[7];
#pragma unroll
      for (int k = 1; k < 7; ++k) {
        const bool h = (hm >> k) & 1u;
        const double w = (k == 3 || k == 5 || k == 6) ? 0.25 : 0.5;
        ck[k] = h ? w : 1.0;
        dk[k] = h ? w : 0.0;
      }
      // x' = P_H x (faces first: they read the raw values of the edge vertices)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double x0 = xe[a], x1 = xe[3 + a], x2 = xe[6 + a], x4 = xe[12 + a];
        xe[9 + a] = ck[3] * xe[9 + a] + dk[3] * ((x0 + x1) + x2);
        xe[15 + a] = ck[5] * xe[15 + a] + dk[5] * ((x0 + x1) + x4);
        xe[18 + a] = ck[6] * xe[18 + a] + dk[6] * ((x0 + x2) + x4);
        xe[3 + a] = ck[1] * x1 + dk[1] * x0;
        xe[6 + a] = ck[2] * x2 + dk[2] * x0;
        xe[12 + a] = ck[4] * x4 + dk[4] * x0;
      }
      // y = P_H^T K_0 x', one row of K_0 at a time: row (k, a) of z = K_0 x' goes to y_k with
      // weight c_k and to vertex 0 (and, for a face, its two edge vertices) with weight d_k.
      // kzero is 0: each row's address formally depends on the round and on the previous
      // row's first entry, so the compiler neither hoists the (loop-invariant) row loads out
      // of the record loop nor issues them all at once (576 values, far more than the scalar
      // registers would hold): the rows are read one after the other through the scalar cache.
      const const_ptr<double> Kr0 = (const_ptr<double>)(K0 + size_t(kzero * __builtin_amdgcn_readfirstlane(base)));
      const_ptr<double> Kr = Kr0;
      double yv[ELAS_ROWS];
#pragma unroll
      for (int r = 0; r < ELAS_ROWS; ++r) yv[r] = 0.0;
#pragma unroll
      for (int r = 0; r < ELAS_ROWS; ++r) {
        double kr[ELAS_ROWS];
#pragma unroll
        for (int q = 0; q < ELAS_ROWS; ++q) kr[q] = Kr[q];
        Kr += ELAS_ROWS + size_t(kzero * uint32_t(__builtin_amdgcn_readfirstlane(int(__double_as_longlong(kr[0])))));
        double zr = kr[0] * xe[0];
#pragma unroll
        for (int q = 1; q < ELAS_ROWS; ++q) zr += kr[q] * xe[q];
        const int k = r / 3, a = r % 3;
        if (k == 0 || k == 7) {
          yv[r] += zr;
        } else {
          yv[r] += ck[k] * zr;
          yv[a] += dk[k] * zr;
          if (k == 3 || k == 5 || k == 6) {
            const int e1 = k & -k, e2 = k & (k - 1);
            yv[3 * e1 + a] += dk[k] * zr;
            yv[3 * e2 + a] += dk[k] * zr;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (l[k] < own_bytes) {
#pragma unroll
          for (int a = 0; a < 3; ++a)
            lds_add(lds_at<unsigned long long>(l[k] + 8u * a), fx_term(yv[3 * k + a], sa[a], 0.0, 0.0, false));
        }
    }
    base += BLOCK;
    if (base >= rec.n_elem) break;
    if (base + threadIdx.x < rec.n_elem) {
      lp = reinterpret_cast<const uint4*>(lid)[rec.elem_off + base + threadIdx.x];
      mt = meta[rec.elem_off + base + threadIdx.x];
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n_own; i += BLOCK) {
    const uint32_t n = nodes[i];
    const bool b = bdy && bdy[n];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double s = fx_value<double>(lds_at<unsigned long long>(0)[3 * i + a], es);
      y[3 * size_t(n) + a] = b ? x[3 * size_t(n) + a] : (finite ? s : __builtin_nan(""));
    }
  }
}

}  // namespace dgh
