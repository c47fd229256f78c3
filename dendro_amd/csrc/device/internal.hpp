// Internal state shared by the device translation units of libdendrohip: device buffers,
// library state, and the mesh / transfer / solver objects behind the opaque C handles.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <vector>

#include "../host/octree.hpp"
#include "kernels.hpp"
#include "vector_kernels.hpp"

namespace dgh {

#define HIP_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) fail(DGH_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------- device buffers
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    release();
    p = o.p;
    n = o.n;
    o.p = nullptr;
    o.n = 0;
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count) {
      if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
        p = nullptr;
        fail(DGH_ERR_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
      }
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T* h, size_t count) {
    alloc(count);
    if (count) HIP_CHECK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
};

static inline unsigned blocks_for(size_t n, int bs = VBLOCK) { return unsigned((n + bs - 1) / bs); }

// ---------------------------------------------------------------- library state
constexpr int OP_LDS_PER_CU = 160 * 1024;   // LDS of one CU (gfx950)
// Default stencil-row source: the class matrices from this many tiles up (four workgroups per
// CU); on fewer tiles the launch is latency-bound and the fewer passes of per-type rows win.
constexpr size_t CLASS_ROWS_MIN_TILES = 1024;
// Meshes of at least this many elements are tiled by default with at most 8192 elements and
// 1024 threads per tile (1.28 records per element on the 44M- and 21M-element finest levels
// of the 8M and 4M point octrees, 8-9% faster than 2048/4096 x 512), smaller ones with 2048 x 512.
constexpr uint32_t LARGE_MESH_ELEMS = 4000000;
// Largest coarsest level solved with a dense inverse (n^2 doubles on the host and the device).
constexpr int COARSE_DENSE_LIMIT = 40000;

struct State {
  bool ready = false;
  int device = 0;
  int max_lds = 65536;            // dynamic LDS bytes a workgroup may use
  DevBuf<double> K64, W64;        // child-0 stencil triangles [18][72]; transfer weights [144][27][8]
  DevBuf<float> K32, W32;
  DevBuf<double> KF64;            // child-0 full stencil rows [18][128] (Laplacian, mass)
  DevBuf<float> KF32;
  std::vector<double> hKF;        // host copy of KF64
  DevBuf<double> SF;              // shape functions at the 3x3x3 Gauss nodes [144][8][27]
  double lrow[NSTENCIL] = {}, mrow[NSTENCIL] = {};   // max row sums of |L|, |M| per type
  rocblas_handle blas = nullptr;
  // device address of the file-scope scratch array of dgh_matvec_dot_noalloc, per device id
  // (looked up by dgh_init with that device current)
  static constexpr int MAX_DEVICES = 64;
  double* noalloc_part[MAX_DEVICES] = {};
};
extern State g;

inline void require_init() {
  if (!g.ready) fail(DGH_ERR_STATE, "dgh_init has not been called");
}

template <class T> const T* Ktab();
template <class T> const T* Wtab();
template <> inline const double* Ktab<double>() { return g.K64.p; }
template <> inline const double* Wtab<double>() { return g.W64.p; }
template <> inline const float* Ktab<float>() { return g.K32.p; }
template <> inline const float* Wtab<float>() { return g.W32.p; }
template <class T> const T* KFtab();
template <> inline const double* KFtab<double>() { return g.KF64.p; }
template <> inline const float* KFtab<float>() { return g.KF32.p; }

}  // namespace dgh

namespace dgh {

// One owner-computes tiling of a mesh (tiles.cpp) and its device copies.
struct DevTiles {
  TilePlan plan;
  uint32_t ys_len = 0;          // owned-node slots per tile (>= max_own, even; batched kernels)
  uint32_t lds64 = 0, lds32 = 0;  // single right-hand-side kernel: LDS bytes of a launch
  int block = 0;                  // single right-hand-side kernel: block size of the plan
  DevBuf<TileRec> tiles;
  DevBuf<uint16_t> lid;         // records: 8 vertex ids with the class in spare bits (pack_records)
  DevBuf<uint32_t> node_tab;
  DevBuf<double> rk64, rm64;    // per-record element factors (kappa h / 2, mass h^3 / 8)
  DevBuf<float> rk32, rm32;
  DevBuf<int2> texp;            // per tile: exponent bound of the contributions, log2 bound of the terms per node
  // uniform coefficients: the element matrices fac1 L + fac2 M of every record class
  // (child-0 type | level << 5), [NCLASS][64], rows of the classes the tiling holds
  DevBuf<double> kc64;
  DevBuf<float> kc32;
};

}  // namespace dgh

namespace dgh {
// Elasticity operator plan of a mesh (elas_api.hip): owner-computes tiles with 24-byte node
// slots, records' meta (child-0 hanging mask, child, level), per-tile fixed-point bounds and
// the unit element's 24x24 matrix of the current Lame parameters.
struct ElasState {
  bool ready = false;
  double mu = 1.0, lam = 1.0;
  TilePlan plan;
  DevBuf<TileRec> tiles;
  DevBuf<uint16_t> lid, meta;   // meta: child-0 hanging mask | child << 8 | level << 11
  DevBuf<uint32_t> node_tab;
  DevBuf<int2> texp;
  DevBuf<double> ke;            // K_0 [24][24]
};
}  // namespace dgh

using namespace dgh;

struct dgh_mesh_s {
  uint32_t E = 0, N = 0, maxD = 0;
  double hFac = 1.0;
  // element arrays in the caller's (Morton) order: transfers and coarse assembly
  std::vector<uint32_t> h_nl;
  std::vector<uint8_t> h_child, h_hn, h_lev, h_cs;
  std::vector<double> h_kappa, h_mass;
  DevBuf<uint32_t> nl;          // [E][8]
  DevBuf<uint8_t> cs;           // [E] stencil index 18c+t
  double lap_fac = 1.0, mass_fac = 1.0;
  int lane_rows = 0;            // stencil-row source of the operator kernels (DGH_ROWS_*)
  bool rows_auto = true;        // choose it at each re-tiling (none set explicitly)
  // owner-computes tilings: single right-hand side, and batched right-hand sides (smaller
  // tiles: the batched kernel stages several values per node in LDS)
  DevTiles op;
  int block = 512;              // threads per operator workgroup (256, 512 or 1024)
  DevTiles mop;
  uint32_t mtile = 0;           // batched tile size (0: default)
  int mvariant = 0;             // batched FP32 products: 0 vector ALUs, 1 matrix cores
  uint64_t gen = 0;             // bumped whenever device tilings or coefficients are rebuilt
  uint64_t coef_gen = 0;        // bumped by dgh_mesh_set_coefficients
  uint64_t bdy_gen = 0;         // bumped by dgh_mesh_set_boundary
  DevBuf<uint8_t> bdy;          // [N] boundary flags of the Dirichlet operators (empty: none)
  std::vector<uint8_t> h_bdy;   // host copy (Dirichlet coarsest matrix)
  DevBuf<double> dotp;          // per-tile partials of dgh_matvec_dot
  DevBuf<int4> geo;             // [E] {anchor x, y, z, level}: element points (dgh_mesh_set_geometry)
  DevBuf<uint8_t> hn;           // [E] hanging masks (nodal values)
  ElasState elas;               // dgh_mesh_set_elasticity
};

struct dgh_transfer_s {
  dgh_mesh coarse = nullptr, fine = nullptr;
  DevBuf<uint32_t> celem;       // [N_f] coarse element holding each fine node
  DevBuf<uint8_t> pos;          // [N_f] position code inside it
  DevBuf<uint32_t> rowptr, col; // R = P^T in CSR (rows: coarse nodes)
  DevBuf<float> w;              // exact dyadic weights
  DevBuf<uint32_t> ent;         // packed (32 w - 1) << 27 | col, when every entry fits
  std::vector<uint32_t> h_celem; // host copies of celem and pos (Jacobian3 coarsest assembly)
  std::vector<uint8_t> h_pos;
  DevBuf<double> j3v64, j3w64;  // fine-level scratch of dgh_matvec_jac3
  DevBuf<float> j3v32, j3w32;
};

namespace dgh {

template <class T>
struct LevelVecs {
  DevBuf<T> invd, b, xa, xb, r;
  DevBuf<T> t;                  // Jacobian3: scratch of the next coarser level's operator
};

}  // namespace dgh

namespace dgh {

// Learned damping of one level (element features, node CSR of (e, k) slots, caps, current
// logits and weights, training samples) and the network state of a solver.
struct LearnLevel {
  bool ready = false;
  DevBuf<uint8_t> lev;
  DevBuf<float> lke, lkbar, cap, dinv32;   // dinv32: plain 1 / A_nn
  DevBuf<double> dinv64;
  DevBuf<uint32_t> rowptr, slot;
  DevBuf<float> O, dO, theta_n, omega, gomega;
  DevBuf<float> e0;                        // samples [N][S]
  int S = 0;
  int age = 0;                             // training steps since the samples were drawn
};

struct LearnState {
  bool init = false;
  DevBuf<float> theta, m1, m2, grad, gpart;
  DevBuf<uint16_t> wbf;
  int step = 0;
  std::vector<LearnLevel> lv;
  DevBuf<float> t, g, u, es, z;            // objective scratch, sized for the finest level
  DevBuf<double> dots;
  std::vector<int> on;                     // levels whose smoother uses the learned weights
};

}  // namespace dgh

struct dgh_solver_s {
  int nlevels = 0;
  std::vector<dgh_mesh> meshes;
  std::vector<dgh_transfer> transfers;
  dgh_solver_params prm{};
  std::vector<LevelVecs<double>> v64;
  std::vector<LevelVecs<float>> v32;
  int nc = 0;                     // coarsest size
  int ldc = 0;                    // row pitch of the dense inverse
  DevBuf<double> Ainv64;          // dense coarsest inverse (full, row pitch ldc)
  DevBuf<float> Ainv32;
  DevBuf<double> Ap64;            // the same in packed lower SB x SB blocks
  DevBuf<float> Ap32;
  DevBuf<double> Qc;              // partial sums of the packed coarse apply [nbc][npc]
  DevBuf<double> c1;              // Ainv 1 (constant null space)
  int nbc = 0, npc = 0;           // coarse block rows, padded coarse size
  bool coarse_iter = false;       // coarsest level solved by Jacobi-preconditioned CG instead
  DevBuf<double> cx, cr, cz, cp, cw;   // its FP64 vectors (coarsest size)
  DevBuf<double> cS;              // its scalars (CoarseCgScalar)
  double* hcS = nullptr;          // pinned host mirror of cS
  int coarse_its = 0;             // iterations of the last iterative coarse solve
  DevBuf<double> S;               // CG scalars (CgScalar)
  DevBuf<double> part;            // reduction partials [MAXQ][RED_BLOCKS]
  DevBuf<double> part_p;          // partials of the null-space projections
  double* hS = nullptr;           // pinned host mirror of S
  DevBuf<double> x, r, z, p, w;   // outer CG vectors (finest, FP64)
  DevBuf<double> ptile;           // per-tile partials of p.w from the operator kernel
  const void* zres = nullptr;     // V-cycle result buffer on the finest level (z, unprojected)
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;     // one CG iteration
  hipGraphExec_t gexec = nullptr;
  uint64_t graph_gen = 0;         // mesh_generation() when the graphs were captured
  // batched right-hand sides (sized for mk columns on first use)
  int mk = 0;
  std::vector<LevelVecs<float>> m32;   // per level [N_l][mk]: b, xa, xb, r
  DevBuf<double> mx, mr, mz, mp, mw;   // outer CG vectors [N][mk]
  DevBuf<double> mS;                   // per-column CG scalars
  DevBuf<double> mpart, mpart_p;       // column reduction partials
  DevBuf<double> mptile;               // CG mode of the FP64 batched product: [tiles][k] partials of p.w
  DevBuf<double> mdense;               // split partial sums of the batched coarse solve
  double* hmS = nullptr;
  hipGraph_t mgraph = nullptr;
  hipGraphExec_t mgexec = nullptr;
  std::vector<double> om_level;        // Richardson scale per level (1 where learned weights sit in invd)
  bool dirichlet = false;              // every level carries boundary flags: the Dirichlet problem
  std::vector<uint64_t> coef_gen;      // meshes' coef_gen / bdy_gen when the diagonals and the
  std::vector<uint64_t> bdy_gen;       // coarsest inverse were built
  LearnState* learn = nullptr;
  ~dgh_solver_s() {
    delete learn;
    if (mgexec) (void)hipGraphExecDestroy(mgexec);
    if (mgraph) (void)hipGraphDestroy(mgraph);
    if (hmS) (void)hipHostFree(hmS);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (hS) (void)hipHostFree(hS);
    if (hcS) (void)hipHostFree(hcS);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace dgh {
void upload_coefs(dgh_mesh m, DevTiles& d);
// Sum of the generations of a solver's meshes: captured graphs hold their buffers.
inline uint64_t mesh_generation(dgh_solver s) {
  uint64_t g = 0;
  for (dgh_mesh m : s->meshes) g += m->gen;
  return g;
}
void build_devtiles(dgh_mesh m, DevTiles& d, uint32_t T, size_t (*lds)(const DevTiles&), uint32_t budget = 0,
                    int block = 0);
}  // namespace dgh

namespace dgh {
// batched pieces shared by the batched solver and the learned-damping trainer (multi_api.hip)
void ensure_multi(dgh_mesh m);
void ensure_multi_buffers(dgh_solver s, int k);
void multi_op_f32(dgh_mesh m, int mode, const float* x, float* y, const float* b, const float* invd, float om,
                  int k, hipStream_t st);
float* mvcycle(dgh_solver s, int l, int k, const float* b, hipStream_t st, bool x1_ready = false);
void restrict_multi_f32(dgh_transfer t, int k, const float* f, float* c, hipStream_t st);
void prolong_multi_f32(dgh_transfer t, int k, const float* c, const float* f_in, float* f_out, hipStream_t st);
void project_mean_multi_f32(dgh_solver s, uint32_t n, int k, float* v, hipStream_t st);
void column_dots_f32(uint32_t n, int k, const float* a, const float* b, double* part, double* out, hipStream_t st);
void drop_graphs(dgh_solver s);
void refresh_operators(dgh_solver s);
}  // namespace dgh
