// CDNA4 (gfx950) operator and transfer kernels of libdendrohip.
//
// The operator kernel is owner-computes over tiles of Morton-ordered elements (the host
// plan of tiles.cpp): one workgroup per tile stages the tile's x values in LDS, evaluates
// every record (own and halo elements) with the 18 child-0 stencils read through the scalar
// cache, accumulates the rows of the nodes the tile owns with 64-bit fixed-point LDS
// atomics, and stores each owned node exactly once.  There are no global atomics, y needs no
// zeroing, and the integer sums make the result independent of the order of the additions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "host/octree.hpp"

namespace dgh {

constexpr int NXCD = 8;          // XCDs of an MI355X; workgroups are dealt round-robin
constexpr int NSTENCIL = 18;     // child-0 stencil types
constexpr int NCLASS = 1024;     // record classes: child-0 type | level << 5 (levels <= 31)

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
// Per-element coefficients arrive with the factors already formed per record by the host
// (h is a power of two, so the products are exact either way); uniform ones are scaled here.
__device__ __forceinline__ void elem_factors(uint32_t lev, uint32_t maxD, double hFac, double k1,
                                             double k2, double& fac1, double& fac2) {
  const double h = hFac * double(1u << (maxD - lev));
  fac1 = k1 * h / 2.0;
  fac2 = k2 * h * h * h / 8.0;
}

// MATVEC: y = A x;  DIAG: y = diag(A);  RESID: y = b - A x;
// SMOOTH: y = x + omega (invd (b - A x))  (one Richardson + Jacobi sweep, y must not alias x);
// CG: y = A x and the tile's partial sum of x.y (the CG product p.Ap).
// OP_DIRMV: the Dirichlet-masked MatVec (examples/src/backend/omgJac.C:859-1056): x of
// boundary nodes is staged as 0 (their columns drop out) and their rows return x itself.
enum OpMode { OP_MATVEC = 0, OP_DIAG = 1, OP_RESID = 2, OP_SMOOTH = 3, OP_CG = 4, OP_DIRMV = 5 };

template <class T>
struct OpIO {
  const T* x;      // operator input (MATVEC, RESID, SMOOTH, CG)
  T* y;            // output
  const T* b;      // right-hand side (RESID, SMOOTH)
  const T* invd;   // inverse diagonal (SMOOTH)
  T omega;         // Richardson scale (SMOOTH)
  double* part;    // CG: per-tile partial sums of x.y
  const uint8_t* bdy;   // DIRMV, and DIAG if set: boundary flags per node (nonzero = boundary)
};

// Symmetric 8x8 stencils are kept as their upper triangles (36 entries, row-major):
// entry (k, j) with k <= j sits at tri(k, j) = k*8 - k*(k-1)/2 + (j - k).
__host__ __device__ constexpr int tri(int k, int j) {
  return (k <= j) ? (k * 8 - k * (k - 1) / 2 + (j - k)) : (j * 8 - j * (j - 1) / 2 + (k - j));
}
constexpr int KSTRIDE = 72;   // per stencil type: Laplacian triangle then mass triangle

// One record: tile-local ids of its 8 vertices, child-0 type | level << 5, and its element
// factors (k1 = kappa h / 2, k2 = mass h^3 / 8).
struct Rec {
  uint32_t l[8];
  uint32_t mt;
  double k1, k2;
};

// A record on the device is 16 bytes: the 8 tile-local vertex ids in 14 bits each, the
// record's class (child-0 type | level << 5, 10 bits) in the top 2 bits of ids 0..4
// (pack_records).
constexpr uint32_t REC_ID_MASK = 0x3fffu;
constexpr uint32_t REC_MAX_TOUCH = REC_ID_MASK + 1u;
__device__ __forceinline__ uint32_t rec_class(uint4 lp) {
  return ((lp.x >> 14) & 3u) | ((lp.x >> 30) << 2) | (((lp.y >> 14) & 3u) << 4) | ((lp.y >> 30) << 6) |
         (((lp.z >> 14) & 3u) << 8);
}

// kappa/mass: per-record factors formed by the host, or nullptr for uniform coefficients
// (lap_fac, mass_fac scaled by the element size here).  LSC scales the vertex ids (the
// single right-hand-side kernel takes them as LDS byte offsets).  UNI: 1 uniform, 0 per
// record, -1 decided by kappa at run time.
// FAC false: the factors are not formed (class matrices carry them).
template <class TC, bool MASS, uint32_t LSC = 1, int UNI = -1, bool FAC = true>
__device__ __forceinline__ void load_rec(Rec& R, uint32_t g, const uint16_t* __restrict__ lid,
                                         const TC* __restrict__ kappa, const TC* __restrict__ mass,
                                         double lap_fac, double mass_fac, uint32_t maxD, double hFac) {
  const uint4 lp = reinterpret_cast<const uint4*>(lid)[g];
  R.l[0] = (lp.x & REC_ID_MASK) * LSC; R.l[1] = ((lp.x >> 16) & REC_ID_MASK) * LSC;
  R.l[2] = (lp.y & REC_ID_MASK) * LSC; R.l[3] = ((lp.y >> 16) & REC_ID_MASK) * LSC;
  R.l[4] = (lp.z & REC_ID_MASK) * LSC; R.l[5] = ((lp.z >> 16) & REC_ID_MASK) * LSC;
  R.l[6] = (lp.w & REC_ID_MASK) * LSC; R.l[7] = ((lp.w >> 16) & REC_ID_MASK) * LSC;
  R.mt = rec_class(lp);
  if (UNI < 0 ? kappa != nullptr : UNI == 0) {
    R.k1 = double(kappa[g]);
    R.k2 = MASS ? double(mass[g]) : 0.0;
  } else if (FAC) {
    elem_factors(R.mt >> 5, maxD, hFac, lap_fac, mass_fac, R.k1, R.k2);
  }
}

// Full stencil rows for the single right-hand-side kernel: per child-0 type, the Laplacian
// rows [8][8] then the mass rows [8][8].
constexpr int KFSTRIDE = 128;

// Owned-node rows accumulate in fixed point: 64-bit integers for FP64 operators, 32-bit for
// FP32.  FP64: a contribution c enters as the integer round(c 2^es), formed by one FMA with
// 1.5 2^52: for |c 2^es| < 2^51 the bit pattern of fma(c, 2^es, 1.5 2^52) is that of
// 1.5 2^52 plus round(c 2^es), so subtracting the former leaves the integer in two's
// complement; es keeps every term below 2^FX_BITS.  FP32: round-to-nearest conversion of
// c 2^es (float), es keeping every term below 2^(FX32_BITS - v) where 2^v bounds the number of
// terms a node of the tile receives, so the 32-bit sums cannot overflow.  es is chosen per
// tile: the host bounds the element factors times the stencil row sums (texp[t].x: bound
// < 2^texp) and the terms per node (texp[t].y = v), the kernel bounds max |x| over the tile's
// staged values.  Each FP64 term is rounded to 2^-FX_BITS of the tile's bound (output ~1e-14
// relative), each FP32 term to about 2^-24 of it.  Integer LDS additions run faster than
// floating-point ones on scattered addresses (ds_add_u64 about twice ds_add_f64, ds_add_u32
// faster again) and do not depend on the order of the additions, so the operator is
// bit-reproducible.
constexpr int FX_BITS = 50;
constexpr int FX32_BITS = 30;
constexpr double FX_MAGIC = 6755399441055744.0;   // 1.5 * 2^52
constexpr unsigned long long FX_MAGIC_BITS = 0x4338000000000000ull;
constexpr int FX_XMIN = -900;                      // |x| below 2^-900 counts as 2^-900

template <class T>
using fx_acc_t = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type;

// Scale exponent of a tile: every contribution times 2^es stays below the fixed-point bound.
template <class T>
__device__ __forceinline__ int fx_scale(int2 te, int ex) {
  return sizeof(T) == 8 ? FX_BITS - te.x - ex : FX32_BITS - te.y - te.x - ex;
}

// fixed-point term of a + b (a = row sum times the scaled Laplacian factor, b likewise mass)
__device__ __forceinline__ unsigned long long fx_term(double a, double f1s, double b, double f2s, bool mass) {
  const double v = mass ? __builtin_fma(a, f1s, __builtin_fma(b, f2s, FX_MAGIC)) : __builtin_fma(a, f1s, FX_MAGIC);
  return static_cast<unsigned long long>(__double_as_longlong(v)) - FX_MAGIC_BITS;
}
__device__ __forceinline__ unsigned int fx_term(float a, float f1s, float b, float f2s, bool mass) {
  const float v = mass ? __builtin_fmaf(b, f2s, a * f1s) : a * f1s;
  return static_cast<unsigned int>(__float2int_rn(v));
}

template <class T>
__device__ __forceinline__ double fx_value(fx_acc_t<T> s, int es) {
  return sizeof(T) == 8 ? ldexp(double(static_cast<long long>(s)), -es)
                        : ldexp(double(static_cast<int>(static_cast<unsigned int>(s))), -es);
}

// Fixed LDS layout of one operator tile.  tile_op_kernel declares no static LDS, so its
// dynamic LDS starts at address 0 and each region sits at a compile-time byte offset.  The
// sums start at address 0 and the staged x's offset rides in the immediate field of the ds
// instructions: a record corner costs one address computation (its id times the element
// size) for both its x read and its sum.
//   [0, tile_xhi_off)             fixed-point sums of the owned nodes, op_own_cap(BLOCK) entries
//   [tile_xhi_off, + 64)          high word of max |x|, one per wavefront (at most 16)
//   [tile_xs_off, ...)            staged x of the touched nodes, owned first
// The host plan keeps n_own <= op_own_cap(BLOCK) and op_own_cap(BLOCK) + n_touch within the
// block size's node budget.
__host__ __device__ constexpr uint32_t op_own_cap(int block) {
  return block >= 1024 ? 4096u : block >= 512 ? 2048u : 1024u;
}
template <class T, int BLOCK>
__host__ __device__ constexpr uint32_t tile_xhi_off() {
  return uint32_t(sizeof(fx_acc_t<T>)) * op_own_cap(BLOCK);
}
template <class T, int BLOCK>
__host__ __device__ constexpr uint32_t tile_xs_off() {
  return tile_xhi_off<T, BLOCK>() + 64u;
}
// Lane stencils (small levels): a copy of the stencil table follows the staged x of the
// tile, 16-byte aligned.
constexpr uint32_t KF_ELEMS = NSTENCIL * KFSTRIDE;
template <class T>
__host__ __device__ constexpr uint32_t tile_kf_off(uint32_t xs_off, uint32_t n_touch) {
  return (xs_off + uint32_t(sizeof(T)) * n_touch + 15u) & ~15u;
}
template <class T>
__host__ constexpr size_t tile_lds_bytes(int block, uint32_t max_touch, bool lane_rows = false) {
  const uint32_t xs = uint32_t(sizeof(fx_acc_t<T>)) * op_own_cap(block) + 64u;
  return lane_rows ? size_t(tile_kf_off<T>(xs, max_touch)) + sizeof(T) * KF_ELEMS
                   : size_t(xs) + sizeof(T) * size_t(max_touch);
}
template <class U>
using lds_ptr = __attribute__((address_space(3))) U*;
template <class U>
__device__ __forceinline__ lds_ptr<U> lds_at(uint32_t byte) {
  return (lds_ptr<U>)byte;
}
template <class U>
__device__ __forceinline__ void lds_add(lds_ptr<U> p, U v) {
  __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
}

// Batched kernels: evaluates one record for CPL adjacent columns with stencil triangles
// Ls/Ms held in scalar registers (one pass over the 36 triangle entries: each stencil value
// feeds both rows it belongs to) and adds the rows of owned nodes, column q scaled by
// 2^es[q], to the fixed-point sums.  R.l holds the nodes' tile-local ids; node l's values
// of the lane's columns sit at LDS byte address xb + l * RS * sizeof(T) (x) and
// yb + l * RS * sizeof(acc) (sums), one 24-bit multiply-add per address (written out: the
// compiler shares l * stride between the two addresses and adds twice; the strides are at
// most 64, inline constants).
template <class T, bool MASS, int MODE, int CPL, int RS>
__device__ __forceinline__ void eval_rec_tri(const Rec& R, uint32_t n_own, const int* es, const T* Ls,
                                             const T* Ms, uint32_t xb, uint32_t yb) {
  using TA = fx_acc_t<T>;
  const T fac1 = T(R.k1), fac2 = T(R.k2);
  T f1s[CPL], f2s[CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    if constexpr (sizeof(T) == 4) {
      f1s[q] = ldexpf(fac1, es[q]);
      f2s[q] = MASS ? ldexpf(fac2, es[q]) : T(0);
    } else {
      f1s[q] = ldexp(fac1, es[q]);
      f2s[q] = MASS ? ldexp(fac2, es[q]) : T(0);
    }
  }
  T xv[CPL][8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint32_t a;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(a) : "v"(R.l[j]), "n"(int(RS * sizeof(T))), "v"(xb));
#pragma unroll
    for (int q = 0; q < CPL; ++q) xv[q][j] = *lds_at<T>(a + q * uint32_t(sizeof(T)));
  }
  T a[CPL][8], bm[CPL][8];
#pragma unroll
  for (int q = 0; q < CPL; ++q)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[q][k] = T(0);
      bm[q][k] = T(0);
    }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      a[q][k] += Ls[tri(k, k)] * xv[q][k];
      if (MASS) bm[q][k] += Ms[tri(k, k)] * xv[q][k];
    }
#pragma unroll
    for (int j = k + 1; j < 8; ++j) {
      const T l = Ls[tri(k, j)];
      const T m = MASS ? Ms[tri(k, j)] : T(0);
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        a[q][k] += l * xv[q][j];
        a[q][j] += l * xv[q][k];
        if (MASS) {
          bm[q][k] += m * xv[q][j];
          bm[q][j] += m * xv[q][k];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (R.l[k] < n_own) {
      uint32_t ya;
      asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(ya) : "v"(R.l[k]), "n"(int(RS * sizeof(TA))), "v"(yb));
#pragma unroll
      for (int q = 0; q < CPL; ++q)
        lds_add(lds_at<TA>(ya + q * uint32_t(sizeof(TA))), fx_term(a[q][k], f1s[q], bm[q][k], f2s[q], MASS));
    }
}

// Evaluates the rows of the owned vertices of one record and adds them, scaled by 2^es, to
// the fixed-point sums.  Stencil rows are read through the scalar cache, one row at a time
// (16 or 32 SGPRs live): the wavefront walks over the distinct stencil types among its
// lanes (one in class-pure wavefronts, two at the class boundaries of a tile) with a
// wave-uniform row pointer.  The compiler computes a row only when some lane owns it.
// R.l holds LDS byte offsets (ids times sizeof(T)); own_bytes = n_own sizeof(T).
// COMB (uniform coefficients): KF holds the element matrix fac1 L + fac2 M of each record
// class (type | level << 5, [NCLASS][64]); the walk is over classes and a row costs 8
// products instead of 16 plus the two factor products.
template <class T, bool MASS, int MODE, uint32_t XS, bool COMB = false>
__device__ __forceinline__ void eval_rows(const Rec& R, uint32_t own_bytes, int es, const T* __restrict__ KF) {
  if (COMB) {
    const T sc = T(ldexp(1.0, es));
    T xv[8];
    if (MODE != OP_DIAG) {
#pragma unroll
      for (int j = 0; j < 8; ++j) xv[j] = *lds_at<T>(XS + R.l[j]);
    }
    const uint32_t cl = R.mt;
    bool done = false;
    for (;;) {
      const uint64_t pend = __ballot(!done);
      if (pend == 0) break;
      const uint32_t cu = __builtin_amdgcn_readlane(cl, __builtin_ctzll(pend));
      const bool mine = !done && cl == cu;
      const T* Kt = KF + cu * 64u;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        T kr[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) kr[j] = Kt[k * 8 + j];
        if (mine && R.l[k] < own_bytes) {
          T a;
          if (MODE == OP_DIAG) {
            a = kr[k];
          } else {
            a = kr[0] * xv[0];
#pragma unroll
            for (int j = 1; j < 8; ++j) a += kr[j] * xv[j];
          }
          lds_add(lds_at<fx_acc_t<T>>(R.l[k]), fx_term(a, sc, T(0), T(0), false));
        }
      }
      done = done || mine;
    }
    return;
  }
  const T fac1 = T(R.k1), fac2 = T(R.k2);
  const T f1s = T(ldexp(double(fac1), es));
  const T f2s = MASS ? T(ldexp(double(fac2), es)) : T(0);
  T xv[8];
  if (MODE != OP_DIAG) {
#pragma unroll
    for (int j = 0; j < 8; ++j) xv[j] = *lds_at<T>(XS + R.l[j]);
  }
  const uint32_t tp = R.mt & 31u;
  bool done = false;
  for (;;) {
    const uint64_t pend = __ballot(!done);
    if (pend == 0) break;
    const uint32_t tu = __builtin_amdgcn_readlane(tp, __builtin_ctzll(pend));
    const bool mine = !done && tp == tu;
    const T* Kt = KF + tu * KFSTRIDE;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      T lr[8], mr[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        lr[j] = Kt[k * 8 + j];
        if (MASS) mr[j] = Kt[64 + k * 8 + j];
      }
      if (mine && R.l[k] < own_bytes) {
        T a, b = T(0);
        if (MODE == OP_DIAG) {
          a = lr[k];
          if (MASS) b = mr[k];
        } else {
          a = lr[0] * xv[0];
#pragma unroll
          for (int j = 1; j < 8; ++j) a += lr[j] * xv[j];
          if (MASS) {
            b = mr[0] * xv[0];
#pragma unroll
            for (int j = 1; j < 8; ++j) b += mr[j] * xv[j];
          }
        }
        lds_add(lds_at<fx_acc_t<T>>(R.l[k]), fx_term(a, f1s, b, f2s, MASS));
      }
    }
    done = done || mine;
  }
}

// Lane-stencil variant for small levels: each lane reads its own type's rows from the LDS
// copy of the stencil table at kf (byte offset; the row offsets ride in the ds immediate
// field), so a wavefront whose lanes hold several stencil types makes one pass instead of
// one per type (small tiles mix many types in a wavefront; each pass is a chain of eight
// scalar row loads).
template <class T, bool MASS, int MODE, uint32_t XS>
__device__ __forceinline__ void eval_rows_lane(const Rec& R, uint32_t own_bytes, int es, uint32_t kf) {
  const T fac1 = T(R.k1), fac2 = T(R.k2);
  const T f1s = T(ldexp(double(fac1), es));
  const T f2s = MASS ? T(ldexp(double(fac2), es)) : T(0);
  T xv[8];
  if (MODE != OP_DIAG) {
#pragma unroll
    for (int j = 0; j < 8; ++j) xv[j] = *lds_at<T>(XS + R.l[j]);
  }
  const uint32_t rb = kf + (R.mt & 31u) * uint32_t(KFSTRIDE * sizeof(T));
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (R.l[k] < own_bytes) {
      const lds_ptr<T> lr = lds_at<T>(rb + uint32_t(k * 8 * sizeof(T)));
      const lds_ptr<T> mr = lds_at<T>(rb + uint32_t((64 + k * 8) * sizeof(T)));
      T a, b = T(0);
      if (MODE == OP_DIAG) {
        a = lr[k];
        if (MASS) b = mr[k];
      } else {
        a = lr[0] * xv[0];
#pragma unroll
        for (int j = 1; j < 8; ++j) a += lr[j] * xv[j];
        if (MASS) {
          b = mr[0] * xv[0];
#pragma unroll
          for (int j = 1; j < 8; ++j) b += mr[j] * xv[j];
        }
      }
      lds_add(lds_at<fx_acc_t<T>>(R.l[k]), fx_term(a, f1s, b, f2s, MASS));
    }
  }
}

// x gathers per thread in flight together: 12 (the 44M-element Helmholtz MatVec 507 -> 482 us
// against 8), 8 in the FP32 residual and smoothing kernels that also hold the owned nodes' b
// and invd (12 would take them past the registers of two workgroups per CU)
constexpr int XCHUNK = 12;
constexpr int XCHUNK_PREF = 8;

// KF: [18][128] child-0 full stencil rows of precision T (Laplacian, then mass).
// kappa/mass: per record (nullptr: uniform lap_fac / mass_fac).  MASS: mass term present.
// texp[t]: exponent bound of tile t's contributions per unit |x| and log2 bound of the terms
// per node (see FX_BITS).
// LDS per tile: the fixed layout above (tile_xs_off), fixed-point sums u64 for FP64, u32 for
// FP32; the plan must have been built with op_own_cap(BLOCK).
// UNI: uniform coefficients (kappa and mass unused; lap_fac, mass_fac, maxD and hFac form the
// factors), else per-record factors (lap_fac, mass_fac, maxD and hFac unused).  The two are
// separate instantiations so that neither keeps the other's arguments in scalar registers.
// LANE: per-lane stencil rows from an LDS copy of KF (eval_rows_lane) instead of per-type
// passes through the scalar cache.
// COMB (uniform coefficients, not LANE): KF is the tiling's class matrix table (fac1 L +
// fac2 M per record class, eval_rows' COMB form) and the records' factors are not formed.
template <class T, class TC, bool MASS, int MODE, int BLOCK, bool UNI = false, bool LANE = false, bool COMB = false>
__global__ void __launch_bounds__(BLOCK)
tile_op_kernel(const TileRec* __restrict__ tiles, int ntiles, const uint16_t* __restrict__ lid,
               const TC* __restrict__ kappa,
               const TC* __restrict__ mass, double lap_fac, double mass_fac, uint32_t maxD,
               double hFac, const T* __restrict__ KF, const uint32_t* __restrict__ node_tab,
               const int2* __restrict__ texp, OpIO<T> io) {
  constexpr int NW = BLOCK / 64;
  static_assert(!COMB || (UNI && !LANE), "class matrices need uniform coefficients and scalar rows");
  const T* __restrict__ x = io.x;
  static_assert(NW <= 16, "one max |x| word per wavefront in 64 LDS bytes");
  static_assert(sizeof(fx_acc_t<T>) == sizeof(T), "sums and staged x share the byte offsets of a record");
  constexpr uint32_t XS = tile_xs_off<T, BLOCK>();
  static_assert(XS <= 65535u, "staged x must start within the ds immediate offset range");
  const lds_ptr<fx_acc_t<T>> acc = lds_at<fx_acc_t<T>>(0);
  const lds_ptr<T> xs = lds_at<T>(XS);
  const lds_ptr<uint32_t> s_xhi = lds_at<uint32_t>(tile_xhi_off<T, BLOCK>());
  const int t = xcd_tile(blockIdx.x, ntiles);
  if (t >= ntiles) return;
  const TileRec rec = tiles[t];
  const uint32_t* nodes = node_tab + rec.node_off;
  const uint32_t n_own = rec.n_own;
  // first record: its loads are issued before the x gather so both chains overlap
  Rec R;
  if (threadIdx.x < rec.n_elem)
    load_rec<TC, MASS, sizeof(T), UNI ? 1 : 0, !COMB>(R, rec.elem_off + threadIdx.x, lid, kappa, mass, lap_fac,
                                                      mass_fac, maxD, hFac);
  const uint32_t kf = tile_kf_off<T>(XS, rec.n_touch);
  if (LANE) {   // the stencil table, copied while the gather runs (visible after the barrier)
    for (uint32_t i = threadIdx.x; i < KF_ELEMS; i += BLOCK) lds_at<T>(kf)[i] = KF[i];
  }
  uint32_t xhi = 0;   // high word of max |x| over the values this thread staged
  // FP32 residual and smoothing with up to 512 threads: the owned nodes' ids, b and invd are
  // loaded with the first gather batch, so the store phase waits on no global load (tiles of
  // at most XC * BLOCK owned nodes; larger ones reload them).  In FP64, and with 1024-thread
  // tiles (two workgroups per CU need at most 64 VGPRs), the held values cost more than the
  // reloads (8M-point octree, 1024-thread finest tiles: mixed PCG 36.4 ms holding 8 per
  // thread, 47.9 ms holding 12, 36.0 ms holding none).
  constexpr bool PREF = sizeof(T) == 4 && (MODE == OP_RESID || MODE == OP_SMOOTH) && BLOCK < 1024;
  constexpr int XC = PREF ? XCHUNK_PREF : XCHUNK;
  const bool pref = PREF && n_own <= uint32_t(XC * BLOCK);
  uint32_t own_n[XC];
  T own_b[XC], own_d[XC];
  // the other kernels keep the owned nodes' ids from the gather for the store phase (tiles
  // whose touched nodes fit one gather batch; larger ones reload them), except the Dirichlet
  // product, whose boundary flags already take it past 64 VGPRs with them
  constexpr bool KEEP = !PREF && MODE != OP_DIAG && MODE != OP_DIRMV;
  const bool keep = KEEP && rec.n_touch <= uint32_t(XC * BLOCK);
  if (MODE != OP_DIAG) {
    for (uint32_t base = 0; base < rec.n_touch; base += XC * BLOCK) {
      uint32_t nid[XC];
      T v[XC];
#pragma unroll
      for (int q = 0; q < XC; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        nid[q] = (i < rec.n_touch) ? nodes[i] : 0u;
      }
      if (KEEP && base == 0) {
#pragma unroll
        for (int q = 0; q < XC; ++q) own_n[q] = nid[q];
      }
      if (PREF && base == 0) {
#pragma unroll
        for (int q = 0; q < XC; ++q) {
          const uint32_t i = threadIdx.x + q * BLOCK;
          own_n[q] = nid[q];
          own_b[q] = (pref && i < n_own) ? io.b[nid[q]] : T(0);
          own_d[q] = (MODE == OP_SMOOTH && pref && i < n_own) ? io.invd[nid[q]] : T(0);
        }
      }
#pragma unroll
      for (int q = 0; q < XC; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        v[q] = (i < rec.n_touch) ? x[nid[q]] : T(0);
      }
      if (MODE == OP_DIRMV) {   // flags loaded beside x (both wait on the node ids only)
        uint8_t bf[XC];
#pragma unroll
        for (int q = 0; q < XC; ++q) {
          const uint32_t i = base + threadIdx.x + q * BLOCK;
          bf[q] = (i < rec.n_touch) ? io.bdy[nid[q]] : uint8_t(0);
        }
#pragma unroll
        for (int q = 0; q < XC; ++q) v[q] = bf[q] ? T(0) : v[q];
      }
#pragma unroll
      for (int q = 0; q < XC; ++q) {
        const uint32_t i = base + threadIdx.x + q * BLOCK;
        if (i < rec.n_touch) xs[i] = v[q];
        xhi = max(xhi, uint32_t(uint64_t(__double_as_longlong(double(v[q]))) >> 32) & 0x7fffffffu);
      }
    }
  }
  for (uint32_t i = threadIdx.x; i < n_own; i += BLOCK) acc[i] = 0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) xhi = max(xhi, uint32_t(__shfl_xor(int(xhi), off)));
  if ((threadIdx.x & 63) == 0) s_xhi[threadIdx.x >> 6] = xhi;
  __syncthreads();
#pragma unroll
  for (int w = 0; w < NW; ++w) xhi = max(xhi, s_xhi[w]);
  // max |x| < 2^ex; a non-finite x makes every owned output of the tile NaN
  const bool finite = (xhi >> 20) < 0x7ffu;
  const int ex = MODE == OP_DIAG ? 1 : max(int(xhi >> 20) - 1022, FX_XMIN);
  const int es = fx_scale<T>(texp[t], ex);
  for (uint32_t base = 0;;) {
    if (base + threadIdx.x < rec.n_elem) {
      if (LANE)
        eval_rows_lane<T, MASS, MODE, XS>(R, n_own * uint32_t(sizeof(T)), es, kf);
      else
        eval_rows<T, MASS, MODE, XS, COMB>(R, n_own * uint32_t(sizeof(T)), es, KF);
    }
    base += BLOCK;
    if (base >= rec.n_elem) break;
    if (base + threadIdx.x < rec.n_elem) {   // wave-uniform bases, lane offsets fixed
      const uint32_t r0 = rec.elem_off + base;
      load_rec<TC, MASS, sizeof(T), UNI ? 1 : 0, !COMB>(R, threadIdx.x, lid + size_t(r0) * 8,
                                                 UNI ? kappa : kappa + r0, UNI || !MASS ? mass : mass + r0, lap_fac,
                                                 mass_fac, maxD, hFac);
    }
  }
  // FP32 residual and smoothing with kept ids: the owned nodes' b and invd are loaded before
  // the barrier, in flight while the other wavefronts finish their records (8M-point octree,
  // mixed PCG 35.2 -> 34.7 ms; in FP64 the 24 held values take the 1024-thread kernel past
  // 64 VGPRs: FP64 PCG 43.9 -> 51.5 ms)
  constexpr bool BPRE = KEEP && sizeof(T) == 4 && (MODE == OP_RESID || MODE == OP_SMOOTH);
  if (BPRE && keep) {
#pragma unroll
    for (int q = 0; q < XC; ++q) {
      const uint32_t i = threadIdx.x + q * BLOCK;
      own_b[q] = i < n_own ? io.b[own_n[q]] : T(0);
      own_d[q] = (MODE == OP_SMOOTH && i < n_own) ? io.invd[own_n[q]] : T(0);
    }
  }
  __syncthreads();
  if (PREF && pref) {
#pragma unroll
    for (int q = 0; q < XC; ++q) {
      const uint32_t i = threadIdx.x + q * BLOCK;
      if (i < n_own) {
        const T v = finite ? T(fx_value<T>(acc[i], es)) : T(__builtin_nan(""));
        if (MODE == OP_RESID)
          io.y[own_n[q]] = own_b[q] - v;
        else
          io.y[own_n[q]] = xs[i] + io.omega * ((own_b[q] - v) * own_d[q]);
      }
    }
    return;
  }
  double dot = 0.0;
  for (uint32_t i = threadIdx.x, q = 0; i < n_own; i += BLOCK, ++q) {
    uint32_t n = 0;
    if (KEEP && keep) {
#pragma unroll
      for (int u = 0; u < XC; ++u) n = (uint32_t(u) == q) ? own_n[u] : n;
    } else {
      n = nodes[i];
    }
    const double s = fx_value<T>(acc[i], es);
    const T v = finite ? T(s) : T(__builtin_nan(""));
    if (MODE == OP_MATVEC) {
      io.y[n] = v;
    } else if (MODE == OP_DIAG) {   // with boundary flags, the Dirichlet diagonal: 1 there
      io.y[n] = (io.bdy && io.bdy[n]) ? T(1) : v;
    } else if (MODE == OP_DIRMV) {
      io.y[n] = io.bdy[n] ? x[n] : v;
    } else if (MODE == OP_RESID || MODE == OP_SMOOTH) {
      T bn, dn;
      if (BPRE && keep) {
        bn = own_b[0];
        dn = own_d[0];
#pragma unroll
        for (int u = 1; u < XC; ++u) {
          bn = (uint32_t(u) == q) ? own_b[u] : bn;
          dn = (uint32_t(u) == q) ? own_d[u] : dn;
        }
      } else {
        bn = io.b[n];
        dn = MODE == OP_SMOOTH ? io.invd[n] : T(0);
      }
      if (MODE == OP_RESID)
        io.y[n] = bn - v;
      else
        io.y[n] = xs[i] + io.omega * ((bn - v) * dn);   // owned nodes lead the staged x
    } else {
      io.y[n] = v;
      dot += double(xs[i]) * double(v);
    }
  }
  if (MODE == OP_CG) {   // tile partial of p'.y, summed in a fixed order
    __syncthreads();
    const lds_ptr<double> red = lds_at<double>(0);   // over the sums, dead past the barrier
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
      for (int w = 0; w < NW; ++w) tot += red[w];
      io.part[t] = tot;
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
               T* __restrict__ f_out, const uint8_t* __restrict__ cbdy = nullptr,
               const uint8_t* __restrict__ fbdy = nullptr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  const uint32_t ce = celem[j];
  const T* w = W + (size_t(cs[ce]) * 27 + pos[j]) * 8;
  const uint4 a = reinterpret_cast<const uint4*>(cnl)[2 * ce];
  const uint4 b = reinterpret_cast<const uint4*>(cnl)[2 * ce + 1];
  const uint32_t idx[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  // the 8 coarse values are loaded whatever the weights (they sit in L2): the loads then wait
  // on the vertex ids only, not on the weights as well; zero weights still add nothing
  T cv[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cv[k] = c[idx[k]];
  if (cbdy) {   // Dirichlet: no entry from a boundary coarse node
#pragma unroll
    for (int k = 0; k < 8; ++k) cv[k] = cbdy[idx[k]] ? T(0) : cv[k];
  }
  T acc = T(0);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T wk = w[k];
    acc += (wk != T(0)) ? wk * cv[k] : T(0);
  }
  if (fbdy && fbdy[j]) acc = T(0);   // Dirichlet: no entry into a boundary fine node
  f_out[j] = (f_in ? f_in[j] : T(0)) + acc;
}

// c[i] = sum_p w_p f[col_p] over row i of R = P^T (CSR built on the host; gather, no
// atomics).  RLANES lanes share a row (rows hold about 4-30 entries) and combine their
// partial sums with shuffles; the sum order is fixed.
constexpr int RLANES = 4;
// With x1 given, also the coarse level's first smoothing sweep from a zero guess:
// x1 = omega (c * invd), as smooth_first_kernel forms it.
// PACKED: one 32-bit word per entry, (32 w - 1) << 27 | fine node (every weight is a
// multiple of 1/32 in (0, 1]); otherwise col and w arrays.
constexpr int RP_SHIFT = 27;
constexpr uint32_t RP_COL = (1u << RP_SHIFT) - 1u;
template <class T, bool PACKED = false>
__global__ void __launch_bounds__(256)
restrict_kernel(uint32_t nc, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                const float* __restrict__ w, const T* __restrict__ f, T* __restrict__ c, T omega = T(0),
                const T* __restrict__ invd = nullptr, T* __restrict__ x1 = nullptr,
                const uint8_t* __restrict__ cbdy = nullptr, const uint8_t* __restrict__ fbdy = nullptr) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = gid / RLANES, lane = gid % RLANES;
  T acc = T(0);
  if (i < nc) {
    const uint32_t p1 = rowptr[i + 1];
    // RU entries of a lane per step: their column loads, then their f loads, go out together
    // (the sum still runs over the lane's entries in order)
    constexpr int RU = 8;
    for (uint32_t p0 = rowptr[i] + lane; p0 < p1; p0 += RU * RLANES) {
      uint32_t e[RU];
      T fv[RU], wv[RU];
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const uint32_t p = p0 + u * RLANES;
        e[u] = p < p1 ? col[p] : 0u;
        wv[u] = (!PACKED && p < p1) ? T(w[p]) : T(0);
      }
#pragma unroll
      for (int u = 0; u < RU; ++u) fv[u] = p0 + u * RLANES < p1 ? f[PACKED ? (e[u] & RP_COL) : e[u]] : T(0);
      if (fbdy) {   // Dirichlet: no entry from a boundary fine node
#pragma unroll
        for (int u = 0; u < RU; ++u)
          if (p0 + u * RLANES < p1 && fbdy[PACKED ? (e[u] & RP_COL) : e[u]]) fv[u] = T(0);
      }
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        if (p0 + u * RLANES >= p1) break;
        if (PACKED)
          acc += T(float((e[u] >> RP_SHIFT) + 1u) * 0.03125f) * fv[u];
        else
          acc += wv[u] * fv[u];
      }
    }
  }
#pragma unroll
  for (int off = RLANES / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, RLANES);
  if (i < nc && lane == 0) {
    if (cbdy && cbdy[i]) acc = T(0);   // Dirichlet: no entry into a boundary coarse node
    c[i] = acc;
    if (x1) x1[i] = T(0) + omega * (acc * invd[i]);
  }
}

template <class T>
__global__ void fill_kernel(T* __restrict__ v, T val, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = val;
}

}  // namespace dgh
