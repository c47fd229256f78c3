// CDNA4 (gfx950) operator and transfer kernels of libdendrohip.
//
// The operator kernel is owner-computes over tiles of Morton-ordered elements (the host
// plan of tiles.cpp): one workgroup per tile stages the tile's x values and the 18 child-0
// stencils in LDS, evaluates every record (own and halo elements), accumulates the rows of
// the nodes the tile owns with LDS atomics, and stores each owned node exactly once.  There
// are no global atomics and y needs no zeroing; the result is deterministic up to the order
// of the LDS additions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host/octree.hpp"

namespace dgh {

constexpr int NXCD = 8;          // XCDs of an MI355X; workgroups are dealt round-robin
constexpr int NSTENCIL = 18;     // child-0 stencil types

// Maps the hardware workgroup id to a tile so that each XCD works on one contiguous
// range of tiles: neighbouring tiles share halo nodes, whose x values then meet in one L2.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
  const int per = (ntiles + NXCD - 1) / NXCD;
  return (bid % NXCD) * per + bid / NXCD;
}

template <class T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

// Element factors exactly as the reference forms them (odaJac.C:271-276,
// omgJac.C:1665-1672): h = hFac 2^(maxD - lev), fac1 = kappa h / 2, fac2 = mass h^3 / 8.
template <class T>
__device__ __forceinline__ void elem_factors(uint32_t lev, uint32_t maxD, double hFac, double k1,
                                             double k2, T& fac1, T& fac2) {
  const double h = hFac * double(1u << (maxD - lev));
  fac1 = T(k1 * h / 2.0);
  fac2 = T(k2 * h * h * h / 8.0);
}

// MATVEC: y = A x;  DIAG: y = diag(A);  RESID: y = b - A x;
// SMOOTH: y = x + omega (invd (b - A x))  (one Richardson + Jacobi sweep, y must not alias x)
enum OpMode { OP_MATVEC = 0, OP_DIAG = 1, OP_RESID = 2, OP_SMOOTH = 3 };

template <class T>
struct OpIO {
  const T* x;      // operator input (MATVEC, RESID, SMOOTH)
  T* y;            // output
  const T* b;      // right-hand side (RESID, SMOOTH)
  const T* invd;   // inverse diagonal (SMOOTH)
  T omega;         // Richardson scale (SMOOTH)
};

// Symmetric 8x8 stencils are kept as their upper triangles (36 entries, row-major):
// entry (k, j) with k <= j sits at tri(k, j) = k*8 - k*(k-1)/2 + (j - k).
__host__ __device__ constexpr int tri(int k, int j) {
  return (k <= j) ? (k * 8 - k * (k - 1) / 2 + (j - k)) : (j * 8 - j * (j - 1) / 2 + (k - j));
}
constexpr int KSTRIDE = 72;   // per stencil type: Laplacian triangle then mass triangle

struct Rec {
  uint32_t l[8];
  uint32_t mt;
  double k1, k2;
};

template <class TC, bool MASS>
__device__ __forceinline__ void load_rec(Rec& R, uint32_t g, const uint16_t* __restrict__ lid,
                                         const uint16_t* __restrict__ meta, const TC* __restrict__ kappa,
                                         const TC* __restrict__ mass, double lap_fac, double mass_fac) {
  const uint4 lp = reinterpret_cast<const uint4*>(lid)[g];
  R.l[0] = lp.x & 0xffffu; R.l[1] = lp.x >> 16; R.l[2] = lp.y & 0xffffu; R.l[3] = lp.y >> 16;
  R.l[4] = lp.z & 0xffffu; R.l[5] = lp.z >> 16; R.l[6] = lp.w & 0xffffu; R.l[7] = lp.w >> 16;
  R.mt = meta[g];
  R.k1 = kappa ? double(kappa[g]) : lap_fac;
  R.k2 = kappa ? (MASS ? double(mass[g]) : 0.0) : mass_fac;
}

// Evaluates one record with stencil triangles Ls/Ms (LDS or, when the stencil type is
// uniform across the wavefront, global memory read through the scalar cache); row sums of
// owned nodes are added into ys (accumulator type TA).
template <class T, class TA, bool MASS, int MODE>
__device__ __forceinline__ void eval_rec_with(const Rec& R, uint32_t n_own, uint32_t maxD, double hFac,
                                              const T* Ls, const T* Ms, const T* __restrict__ xs,
                                              TA* __restrict__ ys) {
  T fac1, fac2;
  elem_factors<T>(R.mt >> 5, maxD, hFac, R.k1, R.k2, fac1, fac2);
  if (MODE != OP_DIAG) {
    T xv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) xv[j] = xs[R.l[j]];
    // one pass over the 36 triangle entries: each stencil value is read once and feeds
    // both rows it belongs to; the element factors scale the row sums at the end
    T a[8], bm[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = T(0);
      bm[k] = T(0);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] += Ls[tri(k, k)] * xv[k];
      if (MASS) bm[k] += Ms[tri(k, k)] * xv[k];
#pragma unroll
      for (int j = k + 1; j < 8; ++j) {
        const T l = Ls[tri(k, j)];
        a[k] += l * xv[j];
        a[j] += l * xv[k];
        if (MASS) {
          const T m = Ms[tri(k, j)];
          bm[k] += m * xv[j];
          bm[j] += m * xv[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (R.l[k] < n_own) atomic_add(&ys[R.l[k]], TA(MASS ? fac1 * a[k] + fac2 * bm[k] : fac1 * a[k]));
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (R.l[k] < n_own) atomic_add(&ys[R.l[k]], TA(fac1 * Ls[tri(k, k)] + (MASS ? fac2 * Ms[tri(k, k)] : T(0))));
  }
}

// Every stencil read goes through the scalar cache: the wavefront walks over the distinct
// stencil types among its active lanes (one in class-pure wavefronts, two at the class
// boundaries of a tile) and evaluates the lanes of each type with a wave-uniform pointer.
template <class T, class TA, bool MASS, int MODE>
__device__ __forceinline__ void eval_rec(const Rec& R, uint32_t n_own, uint32_t maxD, double hFac,
                                         const T* __restrict__ sK, const T* __restrict__ K0,
                                         const T* __restrict__ xs, TA* __restrict__ ys) {
  (void)sK;
  const uint32_t tp = R.mt & 31u;
  bool done = false;
  for (;;) {
    const uint64_t pend = __ballot(!done);
    if (pend == 0) break;
    // stencil type of the first pending lane: wave-uniform, so the stencil is read with
    // scalar loads
    const uint32_t tu = __builtin_amdgcn_readlane(tp, __builtin_ctzll(pend));
    const T* Ls = K0 + tu * KSTRIDE;
    T Lr[36], Mr[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) {
      Lr[q] = Ls[q];
      if (MASS) Mr[q] = Ls[36 + q];
    }
    if (!done && tp == tu) {
      eval_rec_with<T, TA, MASS, MODE>(R, n_own, maxD, hFac, Lr, Mr, xs, ys);
      done = true;
    }
  }
}

// Keep a copy of the 18 stencil triangles in LDS for wavefronts with mixed stencil types
// (false: those few wavefronts read the global table through L1).
constexpr bool LDS_STENCIL = false;
constexpr int RCHUNK = 1;   // records per thread whose loads are in flight together
constexpr int XCHUNK = 4;   // x gathers per thread in flight together

// K0: [18][72] child-0 Laplacian then mass upper triangles of precision T.
// kappa/mass: per record (nullptr: uniform lap_fac / mass_fac).  MASS: mass term present.
// Row sums accumulate in FP64 LDS for both precisions (TA = double).
template <class T, class TC, bool MASS, int MODE, int BLOCK, int RCH = RCHUNK, int XCH = XCHUNK, int MINW = 1>
__global__ void __launch_bounds__(BLOCK, MINW)
tile_op_kernel(const TileRec* __restrict__ tiles, int ntiles, const uint16_t* __restrict__ lid,
               const uint16_t* __restrict__ meta, const TC* __restrict__ kappa,
               const TC* __restrict__ mass, double lap_fac, double mass_fac, uint32_t maxD,
               double hFac, const T* __restrict__ K0, const uint32_t* __restrict__ node_tab,
               uint32_t ys_len, OpIO<T> io) {
  using TA = double;
  const T* __restrict__ x = io.x;
  // LDS: owned-node sums [ys_len >= n_own] (TA), stencils [18][72], x of touched nodes
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TA* ys = reinterpret_cast<TA*>(smem_raw);
  T* sK = LDS_STENCIL ? reinterpret_cast<T*>(ys + ys_len) : nullptr;
  T* xs = LDS_STENCIL ? sK + NSTENCIL * KSTRIDE : reinterpret_cast<T*>(ys + ys_len);
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  const TileRec rec = tiles[t];
  const uint32_t* nodes = node_tab + rec.node_off;
  const uint32_t n_own = rec.n_own;
  // first chunk of records: loads issued before the x gather so both chains overlap
  Rec R[RCH];
#pragma unroll
  for (int q = 0; q < RCH; ++q) {
    const uint32_t r = threadIdx.x + q * BLOCK;
    if (r < rec.n_elem) load_rec<TC, MASS>(R[q], rec.elem_off + r, lid, meta, kappa, mass, lap_fac, mass_fac);
  }
  if (LDS_STENCIL)
    for (int i = threadIdx.x; i < NSTENCIL * KSTRIDE; i += BLOCK) sK[i] = K0[i];
  if (MODE != OP_DIAG) {
    for (uint32_t base = 0; base < rec.n_touch; base += XCH * BLOCK) {
      uint32_t nid[XCH];
      T v[XCH];
#pragma unroll
      for (int q = 0; q < XCH; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        nid[q] = (i < rec.n_touch) ? nodes[i] : 0u;
      }
#pragma unroll
      for (int q = 0; q < XCH; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        v[q] = (i < rec.n_touch) ? x[nid[q]] : T(0);
      }
#pragma unroll
      for (int q = 0; q < XCH; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        if (i < rec.n_touch) xs[i] = v[q];
      }
    }
  }
  for (uint32_t i = threadIdx.x; i < n_own; i += BLOCK) ys[i] = TA(0);
  __syncthreads();
  for (uint32_t base = 0;;) {
#pragma unroll
    for (int q = 0; q < RCH; ++q) {
      const uint32_t r = base + threadIdx.x + q * BLOCK;
      if (r < rec.n_elem) eval_rec<T, TA, MASS, MODE>(R[q], n_own, maxD, hFac, sK, K0, xs, ys);
    }
    base += RCH * BLOCK;
    if (base >= rec.n_elem) break;
#pragma unroll
    for (int q = 0; q < RCH; ++q) {
      const uint32_t r = base + threadIdx.x + q * BLOCK;
      if (r < rec.n_elem) load_rec<TC, MASS>(R[q], rec.elem_off + r, lid, meta, kappa, mass, lap_fac, mass_fac);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n_own; i += BLOCK) {
    const uint32_t n = nodes[i];
    const T v = T(ys[i]);
    if (MODE == OP_MATVEC || MODE == OP_DIAG) {
      io.y[n] = v;
    } else if (MODE == OP_RESID) {
      io.y[n] = io.b[n] - v;
    } else {
      const T r = io.b[n] - v;
      io.y[n] = xs[i] + io.omega * (r * io.invd[n]);   // owned nodes lead the staged x
    }
  }
}

// f_out[j] = f_in[j] + sum_k W[s][pos][k] c[cnl[k]] for every fine node j (prolongation
// P = R^T): the coarse shape functions of the coarse element holding node j.
template <class T>
__global__ void __launch_bounds__(256)
prolong_kernel(uint32_t nf, const uint32_t* __restrict__ celem, const uint8_t* __restrict__ pos,
               const uint32_t* __restrict__ cnl, const uint8_t* __restrict__ cs,
               const T* __restrict__ W, const T* __restrict__ c, const T* __restrict__ f_in,
               T* __restrict__ f_out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  const uint32_t ce = celem[j];
  const T* w = W + (size_t(cs[ce]) * 27 + pos[j]) * 8;
  const uint4 a = reinterpret_cast<const uint4*>(cnl)[2 * ce];
  const uint4 b = reinterpret_cast<const uint4*>(cnl)[2 * ce + 1];
  const uint32_t idx[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  T acc = T(0);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T wk = w[k];
    if (wk != T(0)) acc += wk * c[idx[k]];
  }
  f_out[j] = (f_in ? f_in[j] : T(0)) + acc;
}

// c[i] = sum_p w_p f[col_p] over row i of R = P^T (CSR built on the host; gather, no
// atomics).  RLANES lanes share a row (rows hold about 4-30 entries) and combine their
// partial sums with shuffles; the sum order is fixed.
constexpr int RLANES = 8;
template <class T>
__global__ void __launch_bounds__(256)
restrict_kernel(uint32_t nc, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                const float* __restrict__ w, const T* __restrict__ f, T* __restrict__ c) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = gid / RLANES, lane = gid % RLANES;
  T acc = T(0);
  if (i < nc) {
    const uint32_t p1 = rowptr[i + 1];
    for (uint32_t p = rowptr[i] + lane; p < p1; p += RLANES) acc += T(w[p]) * f[col[p]];
  }
#pragma unroll
  for (int off = RLANES / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, RLANES);
  if (i < nc && lane == 0) c[i] = acc;
}

template <class T>
__global__ void fill_kernel(T* __restrict__ v, T val, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = val;
}

}  // namespace dgh
