// Linear elasticity operator (3 dofs per node) of libdendrohip: the reference's
// ElasticityMatMult (examples/src/backend/elasticityJac.C:129-211),
//   y_k = sum_j fac (mu L[k][j] I + (mu + lambda) GD[3k..3k+2][3j..3j+2]) x_j,  fac = h/2,
// with optional Dirichlet flags (boundary rows y = x, boundary columns dropped).
// Two passes, no atomics:
//   1. one thread per element.  Child c's stencils are child 0's of the mirrored hanging type
//      t' with the vertex order k^c and, for grad-div, the signs s_a s_b (s_a = -1 when bit a
//      of c is set: the mirror flips axis a), so the element takes its inputs as
//      xe[3k+a] = s_a x[3 idx[k^c] + a], applies child 0's stencils of type t' and returns
//      s_a out[3k+a] to vertex k^c.  Elements are sorted by t' inside chunks of ELAS_CHUNK
//      Morton-consecutive elements (the chunk's nodes stay in L2), and every wavefront holds
//      one t': the 8x8 Laplacian and 24x24 grad-div rows are wave-uniform loads.  The 24
//      outputs go to ye[i][24] (i = sorted position, vertex order of the element);
//   2. one thread per node: y[3n+d] = sum over the node's (element, vertex) slots of ye, in a
//      fixed order (a CSR built on the host).
#pragma once

#include <hip/hip_runtime.h>

#include "vector_kernels.hpp"

namespace dgh {

constexpr int ELAS_WAVE = 64;
constexpr uint32_t ELAS_CHUNK = 16384;

// seg[w] = {mirrored type t', first sorted position, count} of wavefront w (count <= 64);
// L0 [18][64] and GD0 [18][576]: child 0's stencils
__global__ void __launch_bounds__(256)
elas_elem_kernel(int nseg, const int4* __restrict__ seg, const uint32_t* __restrict__ order,
                 const uint32_t* __restrict__ nl, const uint8_t* __restrict__ lev, const uint8_t* __restrict__ child,
                 uint32_t maxD, double hFac, double mu, double lam, const double* __restrict__ L0,
                 const double* __restrict__ GD0, const uint8_t* __restrict__ bdy, const double* __restrict__ x,
                 double* __restrict__ ye) {
  const int w = int(blockIdx.x) * (256 / ELAS_WAVE) + int(threadIdx.x) / ELAS_WAVE;
  if (w >= nseg) return;
  const int lane = threadIdx.x % ELAS_WAVE;
  const int4 sg = seg[w];
  if (lane >= sg.z) return;
  const int type = __builtin_amdgcn_readfirstlane(sg.x);
  const uint32_t i = uint32_t(sg.y + lane);
  const uint32_t e = order[i];
  const uint32_t c = child[e];
  const uint4 a4 = reinterpret_cast<const uint4*>(nl)[2 * e];
  const uint4 b4 = reinterpret_cast<const uint4*>(nl)[2 * e + 1];
  const uint32_t idx[8] = {a4.x, a4.y, a4.z, a4.w, b4.x, b4.y, b4.z, b4.w};
  const double sg0 = (c & 1u) ? -1.0 : 1.0, sg1 = (c & 2u) ? -1.0 : 1.0, sg2 = (c & 4u) ? -1.0 : 1.0;
  double xe[24];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t v = idx[0];   // idx[k ^ c] without dynamic register indexing
#pragma unroll
    for (int q = 1; q < 8; ++q) v = (uint32_t(q) == (uint32_t(k) ^ c)) ? idx[q] : v;
    const bool keep = !(bdy && bdy[v]);
    xe[3 * k] = keep ? sg0 * x[3 * size_t(v)] : 0.0;
    xe[3 * k + 1] = keep ? sg1 * x[3 * size_t(v) + 1] : 0.0;
    xe[3 * k + 2] = keep ? sg2 * x[3 * size_t(v) + 2] : 0.0;
  }
  const double fac = hFac * double(1u << (maxD - lev[e])) / 2.0;
  const double f1 = mu * fac, f2 = (mu + lam) * fac;
  const double* __restrict__ Lt = L0 + size_t(type) * 64;
  const double* __restrict__ Gt = GD0 + size_t(type) * 576;
  double* __restrict__ out = ye + size_t(i) * 24;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const double l = Lt[k * 8 + j];
      s0 += l * xe[3 * j];
      s1 += l * xe[3 * j + 1];
      s2 += l * xe[3 * j + 2];
    }
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
    for (int q = 0; q < 24; ++q) {
      g0 += Gt[(3 * k) * 24 + q] * xe[q];
      g1 += Gt[(3 * k + 1) * 24 + q] * xe[q];
      g2 += Gt[(3 * k + 2) * 24 + q] * xe[q];
    }
    const uint32_t ko = uint32_t(k) ^ c;   // the element's own vertex of mirrored vertex k
    out[3 * ko] = sg0 * (f1 * s0 + f2 * g0);
    out[3 * ko + 1] = sg1 * (f1 * s1 + f2 * g1);
    out[3 * ko + 2] = sg2 * (f1 * s2 + f2 * g2);
  }
}

// y[3n+d] = bdy[n] ? x[3n+d] : sum over slots p of node n of ye[slot[p] / 8][3 (slot[p] % 8) + d]
__global__ void __launch_bounds__(VBLOCK)
elas_node_kernel(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ slot,
                 const double* __restrict__ ye, const uint8_t* __restrict__ bdy, const double* __restrict__ x,
                 double* __restrict__ y) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (bdy && bdy[v]) {
    y[3 * size_t(v)] = x[3 * size_t(v)];
    y[3 * size_t(v) + 1] = x[3 * size_t(v) + 1];
    y[3 * size_t(v) + 2] = x[3 * size_t(v) + 2];
    return;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t p = rowptr[v]; p < rowptr[v + 1]; ++p) {
    const uint32_t q = slot[p];
    const double* r = ye + size_t(q >> 3) * 24 + 3 * (q & 7);
    s0 += r[0];
    s1 += r[1];
    s2 += r[2];
  }
  y[3 * size_t(v)] = s0;
  y[3 * size_t(v) + 1] = s1;
  y[3 * size_t(v) + 2] = s2;
}

}  // namespace dgh
