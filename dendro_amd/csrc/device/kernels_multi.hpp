// Batched (multi right-hand side) kernels of libdendrohip.
//
// k right-hand sides are stored node-major: X[n * k + j], so the k values of a node are one
// contiguous row and every node gather is a coalesced row segment.  The operator kernel
// keeps the owner-computes tiling of the single right-hand-side kernel; one workgroup
// handles one tile and one slice of right-hand sides, staging the slice of every touched
// node in LDS.  The element products Y_e = K_e X_e ([8x8] x [8 x slice]) run on the vector
// ALUs (tile_multi_kernel, FP32 and FP64) or, as an FP32 variant, on the matrix cores as
// v_mfma_f32_4x4x1_16b_f32 (tile_multi_mfma_kernel: 16 independent 4x4 blocks per
// instruction, 4 elements x 2 row blocks x 2 column blocks); the autotuner picks the FP32
// variant per mesh.  Row sums of owned nodes accumulate in LDS and are stored once per node.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "kernels.hpp"
#include "vector_kernels.hpp"

namespace dgh {

constexpr int KS = 8;            // right-hand sides per workgroup slice (FP32)
constexpr int KS64 = 4;          // right-hand sides per workgroup slice (FP64: same LDS bytes)
constexpr int MAXK = 32;         // right-hand sides per batch
constexpr int MBLOCK = 512;      // threads per batched-operator workgroup (default)

template <class T>
struct MultiIO {
  const T* x;      // [N][k] operator input
  T* y;            // [N][k] output
  const T* b;      // [N][k] right-hand sides (RESID, SMOOTH)
  const T* invd;   // [N] inverse diagonal (SMOOTH)
  T omega;         // Richardson scale (SMOOTH)
  int k;           // right-hand sides per node row
  double* part;    // CG: per tile and column, the owned nodes' partial of x.y ([ntiles][k])
};

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// 16-byte vectors of the gathers and stores: 4 FP32 or 2 FP64 values
template <class T> struct Vec16;
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <class T> using vec16_t = typename Vec16<T>::type;

// Vector-ALU batched operator.  Lane = (element of the round, CPL adjacent right-hand sides
// c .. c + CPL - 1 of the slice): the KSL / CPL lanes of an element share its record, read
// the element's 8 node values of their columns from LDS and accumulate its 8 rows; the
// stencil comes through the scalar cache (per-type wavefront loop), so LDS carries only the
// minimum per element and column: 8 reads and at most 8 additions, while the record, the
// stencil walk and the LDS addresses are shared by the CPL columns of a lane.  Sums
// accumulate in fixed point as in the single right-hand-side kernel (64-bit integers for
// FP64, 32-bit for FP32), with one scale per column of the slice.  Gathers and stores move
// 16-byte vectors of a node's slice.  LDS rows are padded to KSL + 1 values so the lanes of
// different elements spread over the banks.  LDS: owned sums [ys_len][KSL + 1] (fixed
// point), touched x [max_touch][KSL + 1] (T).
template <class T>
__host__ __device__ constexpr int multi_row(int ksl) { return ksl + 1; }

template <class T, bool MASS, int MODE, int KSL, int MB, int CPL>
__global__ void __launch_bounds__(MB)
tile_multi_kernel(const TileRec* __restrict__ tiles, int ntiles, int nslices,
                  const uint16_t* __restrict__ lid,
                  const T* __restrict__ kappa, const T* __restrict__ mass, double lap_fac,
                  double mass_fac, uint32_t maxD, double hFac, const T* __restrict__ K0,
                  const uint32_t* __restrict__ node_tab, const int2* __restrict__ texp, uint32_t ys_len,
                  MultiIO<T> io) {
  static_assert(MODE != OP_DIAG, "the diagonal does not depend on the right-hand side");
  constexpr int LPE = KSL / CPL;              // lanes per element
  static_assert(KSL % CPL == 0 && 64 % LPE == 0, "a wavefront holds whole elements");
  using TA = fx_acc_t<T>;
  constexpr int RS = multi_row<T>(KSL);      // padded LDS row
  constexpr int NW = MB / 64;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ uint32_t s_xhi[NW][KSL];
  TA* ys = reinterpret_cast<TA*>(smem_raw);
  T* xs = reinterpret_cast<T*>(ys + size_t(ys_len) * RS);
  const int np = ntiles * nslices;
  const int pidx = xcd_tile(blockIdx.x, np);   // the slices of a tile stay on one XCD
  if (pidx >= np) return;
  const int t = pidx / nslices, sl = pidx % nslices;
  const TileRec rec = tiles[t];
  const uint32_t* nodes = node_tab + rec.node_off;
  const uint32_t n_own = rec.n_own;
  const size_t k = size_t(io.k);
  // Gathers and stores move 16-byte vectors: LPN lanes cover the KSL columns of a node's
  // slice, VEC columns each (a thread keeps columns cv .. cv + VEC - 1 throughout, since MB
  // is a multiple of LPN).  Indexes past the end are clamped to the last node, so the MGU
  // node-id loads, then the MGU vector loads of a thread go out together with no branch.
  constexpr int VEC = 16 / int(sizeof(T));
  constexpr int LPN = KSL / VEC;
  static_assert(KSL % VEC == 0 && MB % LPN == 0, "whole vectors per node slice");
  using V = vec16_t<T>;
  const int cv = int(threadIdx.x % LPN) * VEC;
  uint32_t xhv[VEC];   // high words of max |x| of this thread's gather columns
#pragma unroll
  for (int e = 0; e < VEC; ++e) xhv[e] = 0u;
  // The owned nodes lead the tile's node list, so when they fit in the first batch of
  // gathers (keep), the thread's store nodes are the node ids of that batch: own_n.
  constexpr int MGU = 4;
  const uint32_t nv = rec.n_touch * LPN;
  const uint32_t nov = n_own * LPN;
  const bool keep = nov <= uint32_t(MGU * MB);
  uint32_t own_n[MGU];
#pragma unroll
  for (int u = 0; u < MGU; ++u) own_n[u] = 0u;
  const T* xsl = io.x + size_t(sl) * KSL + cv;
  for (uint32_t i0 = threadIdx.x; i0 < nv; i0 += MGU * MB) {
    uint32_t nid[MGU];
    V v[MGU];
#pragma unroll
    for (int u = 0; u < MGU; ++u) nid[u] = nodes[min(i0 + u * MB, nv - 1) / LPN];
    if (i0 == threadIdx.x)
#pragma unroll
      for (int u = 0; u < MGU; ++u) own_n[u] = nid[u];
#pragma unroll
    for (int u = 0; u < MGU; ++u) v[u] = *reinterpret_cast<const V*>(xsl + size_t(nid[u]) * k);
#pragma unroll
    for (int u = 0; u < MGU; ++u) {
      const uint32_t i = i0 + u * MB;
      if (i < nv) {
        T* d = xs + (i / LPN) * RS + cv;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          d[e] = v[u][e];
          xhv[e] = max(xhv[e], uint32_t(uint64_t(__double_as_longlong(double(v[u][e]))) >> 32) & 0x7fffffffu);
        }
      }
    }
  }
  for (uint32_t i = threadIdx.x; i < n_own * RS; i += MB) ys[i] = TA(0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int es = lane / LPE, c = (lane % LPE) * CPL;   // element of the round, first column
#pragma unroll
  for (int off = 32; off >= LPN; off >>= 1)
#pragma unroll
    for (int e = 0; e < VEC; ++e) xhv[e] = max(xhv[e], uint32_t(__shfl_xor(int(xhv[e]), off)));
  if (lane < LPN)
#pragma unroll
    for (int e = 0; e < VEC; ++e) s_xhi[wv][cv + e] = xhv[e];
  __syncthreads();
  // per column: max |x| < 2^ex, fixed-point scale 2^esc; a non-finite column gives NaN.
  // (esc: columns c .. c + CPL - 1 of the products)
  int esc[CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    uint32_t xhi = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) xhi = max(xhi, s_xhi[w][c + q]);
    esc[q] = fx_scale<T>(texp[t], max(int(xhi >> 20) - 1022, FX_XMIN));
  }
  // LDS byte addresses of the lane's first column in row 0 of the staged x and of the sums
  const uint32_t xbase = uint32_t(reinterpret_cast<uintptr_t>((lds_ptr<T>)(xs + c)));
  const uint32_t ybase = uint32_t(reinterpret_cast<uintptr_t>((lds_ptr<TA>)(ys + c)));
  constexpr int EPW = 64 / LPE;               // elements per wavefront round
  constexpr int STRIDE = (MB / 64) * EPW;
  // the records of the next round are loaded while the current round is evaluated; the
  // stencil of the wavefront's current type stays in scalar registers across rounds
  Rec Rn;
  uint32_t r0 = uint32_t(wv) * EPW;
  bool hn = r0 + es < rec.n_elem;
  if (hn) load_rec<T, MASS>(Rn, rec.elem_off + r0 + es, lid, kappa, mass, lap_fac, mass_fac, maxD, hFac);
  uint32_t cur = 0xffffffffu;
  T Lr[36], Mr[36];
  for (; r0 < rec.n_elem; r0 += STRIDE) {
    const Rec R = Rn;
    const bool h = hn;
    const uint32_t rn = r0 + STRIDE + es;
    hn = rn < rec.n_elem;
    if (hn) load_rec<T, MASS>(Rn, rec.elem_off + rn, lid, kappa, mass, lap_fac, mass_fac, maxD, hFac);
    const uint32_t tp = R.mt & 31u;
    bool done = !h;
    for (;;) {
      const uint64_t pend = __ballot(!done);
      if (pend == 0) break;
      const uint32_t tu = __builtin_amdgcn_readlane(tp, __builtin_ctzll(pend));
      if (tu != cur) {
        const T* Ls = K0 + tu * KSTRIDE;
#pragma unroll
        for (int q = 0; q < 36; ++q) {
          Lr[q] = Ls[q];
          if (MASS) Mr[q] = Ls[36 + q];
        }
        cur = tu;
      }
      if (!done && tp == tu) {
        eval_rec_tri<T, MASS, MODE, CPL, RS>(R, n_own, esc, Lr, Mr, xbase, ybase);
        done = true;
      }
    }
  }
  // the store's right-hand sides and diagonals of the kept owned nodes: loaded before the
  // barrier, so they are in flight while the other wavefronts finish their records
  constexpr int MSU = MGU;
  const size_t so = size_t(sl) * KSL + cv;
  V bpre[MSU];
  T dpre[MSU];
  if ((MODE == OP_RESID || MODE == OP_SMOOTH) && keep) {
#pragma unroll
    for (int u = 0; u < MSU; ++u) {
      bpre[u] = *reinterpret_cast<const V*>(io.b + size_t(own_n[u]) * k + so);
      dpre[u] = MODE == OP_SMOOTH ? io.invd[own_n[u]] : T(0);
    }
  }
  __syncthreads();
  // scales of the thread's store columns cv + e
  int ess[VEC];
  bool fins[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    uint32_t h = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) h = max(h, s_xhi[w][cv + e]);
    fins[e] = (h >> 20) < 0x7ffu;
    ess[e] = fx_scale<T>(texp[t], max(int(h >> 20) - 1022, FX_XMIN));
  }
  // stores: 16-byte vectors as in the gather, MSU (= MGU, the batch of own_n) per thread
  // with their loads in flight together
  double dot[VEC];   // CG: this thread's partials of x.y for its columns
#pragma unroll
  for (int e = 0; e < VEC; ++e) dot[e] = 0.0;
  for (uint32_t i0 = threadIdx.x; i0 < nov; i0 += MSU * MB) {
    uint32_t nid[MSU];
    V bv[MSU];
    T dv[MSU];
#pragma unroll
    for (int u = 0; u < MSU; ++u) nid[u] = keep ? own_n[u] : nodes[min(i0 + u * MB, nov - 1) / LPN];
#pragma unroll
    for (int u = 0; u < MSU; ++u) {
      if (keep) {
        bv[u] = bpre[u];
        dv[u] = dpre[u];
      } else {
        if (MODE == OP_RESID || MODE == OP_SMOOTH)
          bv[u] = *reinterpret_cast<const V*>(io.b + size_t(nid[u]) * k + so);
        dv[u] = MODE == OP_SMOOTH ? io.invd[nid[u]] : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < MSU; ++u) {
      const uint32_t i = i0 + u * MB;
      if (i >= nov) break;
      const uint32_t nd = i / LPN;
      V out;
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const T v = fins[e] ? T(fx_value<T>(ys[nd * RS + cv + e], ess[e])) : T(__builtin_nan(""));
        if (MODE == OP_MATVEC || MODE == OP_CG) out[e] = v;
        else if (MODE == OP_RESID) out[e] = bv[u][e] - v;
        else out[e] = xs[nd * RS + cv + e] + io.omega * ((bv[u][e] - v) * dv[u]);
        if (MODE == OP_CG) dot[e] += double(xs[nd * RS + cv + e]) * double(v);
      }
      *reinterpret_cast<V*>(io.y + size_t(nid[u]) * k + so) = out;
    }
  }
  if constexpr (MODE == OP_CG) {
    // the tile's partial of x.y per column, summed in a fixed order: over the lanes that
    // share the thread's columns, then over the wavefronts (in LDS over the dead sums)
#pragma unroll
    for (int off = 32; off >= LPN; off >>= 1)
#pragma unroll
      for (int e = 0; e < VEC; ++e) dot[e] += __shfl_xor(dot[e], off);
    __syncthreads();
    double* red = reinterpret_cast<double*>(smem_raw);   // [NW][KSL]
    if (lane < LPN)
#pragma unroll
      for (int e = 0; e < VEC; ++e) red[wv * KSL + cv + e] = dot[e];
    __syncthreads();
    if (threadIdx.x < KSL) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) sum += red[w * KSL + threadIdx.x];
      io.part[size_t(t) * k + size_t(sl) * KSL + threadIdx.x] = sum;
    }
  }
}

// out[j] = sum over tiles of part[t][j] (the CG mode's tile partials), one workgroup per
// column, in a fixed order
static __global__ void __launch_bounds__(256)
sum_tiles_multi_kernel(int ntiles, int k, const double* __restrict__ part, double* __restrict__ out) {
  const int j = blockIdx.x;
  double v = 0.0;
  for (int t = threadIdx.x; t < ntiles; t += 256) v += part[size_t(t) * k + j];
  v = wave_sum(v);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Matrix-core variant (FP32): lane = 4 * block + q, block = 4 * es + 2 * ib + jb: element
// es of the round, row block ib (nodes 4ib..4ib+3), column block jb (right-hand sides
// 4jb..4jb+3).  v_mfma_f32_4x4x1_16b_f32 operands: A = K[4ib + q][kk], B = X[kk][4jb + q];
// result lane 4 * block + q, register r = Y[4ib + r][4jb + q].  The full child-0 stencils
// sit in LDS after the owned sums (FP64): [18][1 + MASS][64].
template <bool MASS>
__host__ __device__ constexpr size_t multi_stencil_len() {
  return size_t(NSTENCIL) * (MASS ? 2 : 1) * 64;
}

template <bool MASS, int MODE, int MB>
__global__ void __launch_bounds__(MB)
tile_multi_mfma_kernel(const TileRec* __restrict__ tiles, int ntiles, int nslices,
                       const uint16_t* __restrict__ lid,
                       const float* __restrict__ kappa, const float* __restrict__ mass, double lap_fac,
                       double mass_fac, uint32_t maxD, double hFac, const float* __restrict__ K0,
                       const uint32_t* __restrict__ node_tab, uint32_t ys_len, MultiIO<float> io) {
  static_assert(MODE != OP_DIAG, "the diagonal does not depend on the right-hand side");
  constexpr int NM = MASS ? 2 : 1;
  constexpr int NV = KS / 4;                  // float4 vectors per node slice
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* ys = reinterpret_cast<double*>(smem_raw);
  float* sK = reinterpret_cast<float*>(ys + size_t(ys_len) * KS);
  float* xs = sK + multi_stencil_len<MASS>();
  const int np = ntiles * nslices;
  const int pidx = xcd_tile(blockIdx.x, np);
  if (pidx >= np) return;
  const int t = pidx / nslices, sl = pidx % nslices;
  const TileRec rec = tiles[t];
  const uint32_t* nodes = node_tab + rec.node_off;
  const uint32_t n_own = rec.n_own;
  const size_t kv = size_t(io.k / 4);
  for (int i = threadIdx.x; i < int(multi_stencil_len<MASS>()); i += MB) {
    const int tp = i / (NM * 64), mm = (i / 64) % NM, kj = i % 64;
    sK[i] = K0[tp * KSTRIDE + mm * 36 + tri(kj >> 3, kj & 7)];
  }
  {
    const float4* X = reinterpret_cast<const float4*>(io.x);
    float4* xsv = reinterpret_cast<float4*>(xs);
    constexpr int MGU = 4;   // gathers per thread in flight together
    const uint32_t nx = rec.n_touch * NV;
    for (uint32_t i0 = threadIdx.x; i0 < nx; i0 += MGU * MB) {
      uint32_t nid[MGU];
      float4 v[MGU];
#pragma unroll
      for (int u = 0; u < MGU; ++u) {
        const uint32_t i = i0 + u * MB;
        nid[u] = i < nx ? nodes[i / NV] : 0u;
      }
#pragma unroll
      for (int u = 0; u < MGU; ++u) {
        const uint32_t i = i0 + u * MB;
        if (i < nx) v[u] = X[size_t(nid[u]) * kv + size_t(sl) * NV + i % NV];
      }
#pragma unroll
      for (int u = 0; u < MGU; ++u) {
        const uint32_t i = i0 + u * MB;
        if (i < nx) xsv[i] = v[u];
      }
    }
  }
  for (uint32_t i = threadIdx.x; i < n_own * KS; i += MB) ys[i] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int NW = MB / 64;
  const int blk = lane >> 2, q = lane & 3;
  const int es = blk >> 2, ib = (blk >> 1) & 1, jb = blk & 1;
  for (uint32_t r0 = uint32_t(wv) * 4; r0 < rec.n_elem; r0 += NW * 4) {
    const uint32_t r = r0 + es;
    const bool act = r < rec.n_elem;
    uint32_t L[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    float a[8], bv[8];
    if (act) {
      Rec R;
      load_rec<float, MASS>(R, rec.elem_off + r, lid, kappa, mass, lap_fac, mass_fac, maxD, hFac);
      const float f1 = float(R.k1), f2 = float(R.k2);
      const float* Lrow = sK + (R.mt & 31u) * (NM * 64) + (4 * ib + q) * 8;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        L[kk] = R.l[kk];
        a[kk] = MASS ? f1 * Lrow[kk] + f2 * Lrow[64 + kk] : f1 * Lrow[kk];
        bv[kk] = xs[R.l[kk] * KS + 4 * jb + q];
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) a[kk] = bv[kk] = 0.f;
    }
    mfma_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a[kk], bv[kk], acc, 0, 0, 0);
    if (act) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const uint32_t nd = L[4 * ib + rr];
        if (nd < n_own) atomic_add(&ys[nd * KS + 4 * jb + q], double(acc[rr]));
      }
    }
  }
  __syncthreads();
  const size_t k = size_t(io.k);
  for (uint32_t i = threadIdx.x; i < n_own * KS; i += MB) {
    const uint32_t nd = i / KS, c = i % KS;
    const size_t gi = size_t(nodes[nd]) * k + size_t(sl) * KS + c;
    const float v = float(ys[i]);
    if (MODE == OP_MATVEC) {
      io.y[gi] = v;
    } else if (MODE == OP_RESID) {
      io.y[gi] = io.b[gi] - v;
    } else {
      const float rr = io.b[gi] - v;
      io.y[gi] = xs[i] + io.omega * (rr * io.invd[nodes[nd]]);
    }
  }
}

// x[n][j] = omega (b[n][j] invd[n])  (first sweep from a zero guess)
template <class T>
__global__ void smooth_first_multi_kernel(uint32_t n, int k, T omega, const T* __restrict__ b,
                                          const T* __restrict__ invd, T* __restrict__ x) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(n) * k) return;
  x[i] = T(0) + omega * (b[i] * invd[i / k]);
}

// c[i][j] = sum_p w_p f[col_p][j]: lanes run over the right-hand sides of a row.
template <class T>
static __global__ void __launch_bounds__(256)
restrict_multi_kernel(uint32_t nc, int k, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                      const float* __restrict__ w, const T* __restrict__ f, T* __restrict__ c) {
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t i = gid / k;
  const int j = int(gid % k);
  if (i >= nc) return;
  T acc = T(0);
  const uint32_t p1 = rowptr[i + 1];
  // four entries per step: their column and weight loads, then their f loads, go out
  // together (the sum still runs over the row in order)
  constexpr int RU = 4;
  for (uint32_t p0 = rowptr[i]; p0 < p1; p0 += RU) {
    uint32_t cc[RU];
    T wv[RU], fv[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const bool in = p0 + u < p1;
      cc[u] = in ? col[p0 + u] : 0u;
      wv[u] = in ? T(w[p0 + u]) : T(0);
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) fv[u] = p0 + u < p1 ? f[size_t(cc[u]) * k + j] : T(0);
#pragma unroll
    for (int u = 0; u < RU; ++u)
      if (p0 + u < p1) acc += wv[u] * fv[u];
  }
  c[i * k + j] = acc;
}

// f_out[j][q] = f_in[j][q] + sum_m W[s][pos_j][m] c[cnl[m]][q]
template <class T>
static __global__ void __launch_bounds__(256)
prolong_multi_kernel(uint32_t nf, int k, const uint32_t* __restrict__ celem, const uint8_t* __restrict__ pos,
                     const uint32_t* __restrict__ cnl, const uint8_t* __restrict__ cs,
                     const T* __restrict__ W, const T* __restrict__ c, const T* __restrict__ f_in,
                     T* __restrict__ f_out) {
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t j = gid / k;
  const int q = int(gid % k);
  if (j >= nf) return;
  const uint32_t ce = celem[j];
  const T* wv = W + (size_t(cs[ce]) * 27 + pos[j]) * 8;
  // the coarse values are loaded whatever the weights: they then wait on the vertex ids only
  const uint4 a = reinterpret_cast<const uint4*>(cnl)[2 * size_t(ce)];
  const uint4 b = reinterpret_cast<const uint4*>(cnl)[2 * size_t(ce) + 1];
  const uint32_t idx[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  T cv[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) cv[m] = c[size_t(idx[m]) * k + q];
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const T wk = wv[m];
    acc += (wk != T(0)) ? wk * cv[m] : T(0);
  }
  f_out[gid] = (f_in ? f_in[gid] : T(0)) + acc;
}

// X = Ainv B for the dense coarsest inverse (row pitch ld) and k right-hand sides ([n][k]),
// split over DM_SPLIT column ranges of Ainv (enough workgroups to fill the chip):
// part[s][i][j] = sum over m in range s of Ainv[i][m] B[m][j] (FP64), then
// dense_multi_sum_kernel adds the ranges in order.  A workgroup computes DM_ROWS rows x 32
// columns of one range, a thread a 4 x 4 block of them (rows 4 tr .. 4 tr + 3, columns
// 4 tc .. 4 tc + 3) from one 16-byte LDS read of each operand per m; Ainv (transposed into
// LDS, rows padded by 4 values) and B pass through LDS in chunks of DM_CH columns of Ainv,
// the next chunk in flight in registers meanwhile.
constexpr int DM_ROWS = 128, DM_CH = 32, DM_SPLIT = 16;
template <class T>
__global__ void __launch_bounds__(256)
dense_multi_kernel(int n, int ld, int k, const T* __restrict__ Ainv, const T* __restrict__ B,
                   double* __restrict__ part) {
  static_assert(DM_ROWS == 4 * 32 && MAXK == 4 * 8, "256 threads of 4 x 4 blocks cover DM_ROWS x MAXK");
  using V = vec16_t<T>;
  static_assert(sizeof(V) == 4 * sizeof(T), "4-value vectors");
  __shared__ __align__(16) T sa[DM_CH][DM_ROWS + 4];
  __shared__ __align__(16) T sb[DM_CH][MAXK];
  const int row0 = blockIdx.x * DM_ROWS;
  const int per = (n + DM_SPLIT - 1) / DM_SPLIT;
  const int mlo = blockIdx.y * per, mhi = min(n, mlo + per);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  // the next chunk's values are loaded into registers while the current chunk is used
  constexpr int NA = DM_ROWS * DM_CH / 256, NB = DM_CH * MAXK / 256;
  T ra[NA], rb[NB];
  auto fetch = [&](int m0) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {   // coalesced along the rows of Ainv
      const int i = threadIdx.x + q * 256, r = i / DM_CH, m = i % DM_CH;
      ra[q] = (row0 + r < n && m0 + m < mhi) ? Ainv[size_t(row0 + r) * ld + m0 + m] : T(0);
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int i = threadIdx.x + q * 256, m = i / MAXK, j = i % MAXK;
      rb[q] = (j < k && m0 + m < mhi) ? B[size_t(m0 + m) * k + j] : T(0);
    }
  };
  if (mlo < mhi) fetch(mlo);
  for (int m0 = mlo; m0 < mhi; m0 += DM_CH) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int i = threadIdx.x + q * 256;
      sa[i % DM_CH][i / DM_CH] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int i = threadIdx.x + q * 256;
      sb[i / MAXK][i % MAXK] = rb[q];
    }
    __syncthreads();
    if (m0 + DM_CH < mhi) fetch(m0 + DM_CH);
#pragma unroll 4
    for (int m = 0; m < DM_CH; ++m) {
      const V a = *reinterpret_cast<const V*>(&sa[m][4 * tr]);
      const V b = *reinterpret_cast<const V*>(&sb[m][4 * tc]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += double(a[u]) * double(b[v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = row0 + 4 * tr + u;
    if (r >= n) break;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (4 * tc + v < k) part[(size_t(blockIdx.y) * n + r) * k + 4 * tc + v] = acc[u][v];
  }
}

template <class T>
__global__ void __launch_bounds__(256)
dense_multi_sum_kernel(int n, int k, const double* __restrict__ part, T* __restrict__ X) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(n) * k) return;
  double s = 0.0;
  for (int q = 0; q < DM_SPLIT; ++q) s += part[size_t(q) * n * k + i];
  X[i] = T(s);
}

// ---------------------------------------------------------------- column reductions
// out[q][j] = sum_n a_q[n][j] * b_q[n][j] (b_q == nullptr: sum of a_q), deterministic:
// stage 1 writes part[(q * RED_BLOCKS_M + block) * k + j], stage 2 sums over blocks in order.
constexpr int RED_BLOCKS_M = 1024;
template <class T>
__global__ void __launch_bounds__(256)
reduce1_multi_kernel(uint32_t n, int k, DotArgs<T> d, double* __restrict__ part) {
  const int j = threadIdx.x % k, rs = threadIdx.x / k, nrs = 256 / k;
  double s[MAXQ] = {0.0, 0.0, 0.0, 0.0};
  const uint32_t per = (n + RED_BLOCKS_M - 1) / RED_BLOCKS_M;
  const uint32_t lo = blockIdx.x * per, hi = min(n, lo + per);
  if (rs < nrs) {
    uint32_t i = lo + rs;
    for (; i + uint32_t(nrs) < hi; i += 2 * nrs) {       // two rows in flight per thread
      const size_t g0 = size_t(i) * k + j, g1 = size_t(i + nrs) * k + j;
#pragma unroll
      for (int q = 0; q < MAXQ; ++q)
        if (d.a[q]) {
          const double v0 = double(d.a[q][g0]) * (d.b[q] ? double(d.b[q][g0]) : 1.0);
          const double v1 = double(d.a[q][g1]) * (d.b[q] ? double(d.b[q][g1]) : 1.0);
          s[q] += v0;
          s[q] += v1;
        }
    }
    for (; i < hi; i += nrs) {
      const size_t g = size_t(i) * k + j;
#pragma unroll
      for (int q = 0; q < MAXQ; ++q)
        if (d.a[q]) s[q] += double(d.a[q][g]) * (d.b[q] ? double(d.b[q][g]) : 1.0);
    }
  }
  __shared__ double red[MAXQ][256];
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  if (threadIdx.x < k) {
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
      double t = 0.0;
      for (int r = 0; r < nrs; ++r) t += red[q][r * k + threadIdx.x];
      part[(size_t(q) * RED_BLOCKS_M + blockIdx.x) * k + threadIdx.x] = t;
    }
  }
}

// out[q * ostride + j]: one workgroup per (q, j), partials summed in a fixed tree order
static __global__ void __launch_bounds__(256)
reduce2_multi_kernel(int nq, int k, const double* __restrict__ part, double* __restrict__ out, int ostride) {
  const int q = blockIdx.x / k, j = blockIdx.x % k;
  if (q >= nq) return;
  double t = 0.0;
  for (int b = threadIdx.x; b < RED_BLOCKS_M; b += 256) t += part[(size_t(q) * RED_BLOCKS_M + b) * k + j];
  t = wave_sum(t);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) out[q * ostride + j] = (red[0] + red[1]) + (red[2] + red[3]);
}

// v[n][j] -= sum[j] / n
template <class T>
__global__ void subtract_mean_multi_kernel(uint32_t n, int k, const double* __restrict__ sum, T* __restrict__ v) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < size_t(n) * k) v[i] = v[i] - T(sum[i % k] / double(n));
}

// ---------------------------------------------------------------- batched CG
// Scalars per right-hand side j: SM[q * MAXK + j] with q a CgScalar, plus SM_DP0 (initial
// norm) and SM_ACT (1 while the column iterates).  A converged column is frozen: its
// step length is zero, so its iterate stays the one that met the test.
enum { SM_DP0 = S_NSCAL, SM_ACT, SM_NSCAL };

static __global__ void cg_p_multi_kernel(uint32_t n, int k, const double* __restrict__ S, const double* __restrict__ z,
                                  double* __restrict__ p) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int j = int(i % k);
  if (i < size_t(n) * k && S[SM_ACT * MAXK + j] != 0.0) p[i] = z[i] + S[S_RATIO * MAXK + j] * p[i];
}

// x += alpha_j p, r -= alpha_j w for the active columns (alpha_j = RZ / PW, stored as ALPHA),
// fused with the start of the batched V-cycle on the finest level: the FP32 right-hand side
// b32 = float(r) and the first sweep from zero x1 = omega (b32 invd), with the arithmetic of
// the conversion and first-sweep kernels.
static __global__ void cg_xr_multi_fused_kernel(uint32_t n, int k, double* __restrict__ S,
                                                const double* __restrict__ p, const double* __restrict__ w,
                                                double* __restrict__ x, double* __restrict__ r,
                                                float* __restrict__ b32, const float* __restrict__ invd,
                                                float omega, float* __restrict__ x1) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int j = int(i % k);
  const bool act = S[SM_ACT * MAXK + j] != 0.0;
  const double a = act ? S[S_RZ * MAXK + j] / S[S_PW * MAXK + j] : 0.0;
  if (i < size_t(k)) S[S_ALPHA * MAXK + j] = a;
  if (i >= size_t(n) * k) return;
  double ri = r[i];
  if (act) {
    x[i] = x[i] + a * p[i];
    ri = ri + (-a) * w[i];
    r[i] = ri;
  }
  const float bf = float(ri);
  b32[i] = bf;
  x1[i] = 0.f + omega * (bf * invd[i / k]);
}

// z = double(v) - mean_j (the column means of v, summed in FP64, in sum[j] * n): the
// conversion of the V-cycle's result and the constant-null-space projection in one pass;
// sum == nullptr: the conversion alone.
static __global__ void convert_sub_mean_multi_kernel(uint32_t n, int k, const float* __restrict__ v,
                                                     const double* __restrict__ sum, double* __restrict__ z) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(n) * k) return;
  const double zi = double(v[i]);
  z[i] = sum ? zi - double(sum[i % k] / double(n)) : zi;
}

static __global__ void cg_scalar_multi_kernel(int k, double* __restrict__ S, int precond_norm, int flexible, double rtol) {
  const int j = threadIdx.x;
  if (j >= k) return;
  if (S[SM_ACT * MAXK + j] == 0.0) return;
  const double dp = sqrt(precond_norm ? S[S_ZZ * MAXK + j] : S[S_RR * MAXK + j]);
  const double ratio = flexible ? (-S[S_ALPHA * MAXK + j] * S[S_ZW * MAXK + j]) / S[S_RZ * MAXK + j]
                                : S[S_ZR * MAXK + j] / S[S_RZ * MAXK + j];
  S[S_DP * MAXK + j] = dp;
  S[S_RATIO * MAXK + j] = ratio;
  S[S_RZ * MAXK + j] = S[S_ZR * MAXK + j];
  if (!(dp > rtol * S[SM_DP0 * MAXK + j])) S[SM_ACT * MAXK + j] = 0.0;
}

static __global__ void cg_init_multi_kernel(int k, double* __restrict__ S, int precond_norm, double rtol) {
  const int j = threadIdx.x;
  if (j >= k) return;
  S[S_RZ * MAXK + j] = S[S_ZR * MAXK + j];
  const double dp = sqrt(precond_norm ? S[S_ZZ * MAXK + j] : S[S_BB * MAXK + j]);
  S[S_DP * MAXK + j] = dp;
  S[SM_DP0 * MAXK + j] = dp;
  S[S_RATIO * MAXK + j] = 0.0;
  S[SM_ACT * MAXK + j] = dp > 0.0 ? 1.0 : 0.0;
  (void)rtol;
}

}  // namespace dgh
