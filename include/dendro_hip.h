/*
 * dendro_hip.h -- C ABI of libdendrohip.so, the MI355X (gfx950) implementation of Dendro's
 * matrix-free octree FEM solve path (oda element-loop MatVec, omg intergrid transfer,
 * damped-Jacobi V-cycle inside PCG).
 *
 * Conventions (all entry points):
 *   - every function returns int: 0 = success, negative = error (DGH_ERR_*);
 *     dgh_last_error() returns the message of the most recent failure on the calling thread;
 *   - plain pointers and sizes only; "_dev" pointers are device memory owned by the caller,
 *     every other pointer is host memory;
 *   - stream arguments are hipStream_t passed as void* (NULL = the null stream); device work is
 *     stream-ordered and the call returns before it completes unless stated otherwise;
 *   - level 0 of a hierarchy is the coarsest and level nlevels-1 the finest, as in the
 *     reference's DAMG array (src/omg/omg.C:1838-1863).
 *
 * The reference interface each entry point replaces is cited beside it (paths relative to the
 * reference tree).
 */
#ifndef DENDRO_HIP_H
#define DENDRO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define DGH_OK 0
#define DGH_ERR_ARG (-1)       /* invalid argument or size                           */
#define DGH_ERR_HIP (-2)       /* HIP runtime error                                  */
#define DGH_ERR_STATE (-3)     /* call out of order (e.g. no dgh_init)               */
#define DGH_ERR_NOMEM (-4)     /* host or device allocation failed                   */
#define DGH_ERR_MESH (-5)      /* mesh invariant violated (invalid hanging mask ...) */
#define DGH_ERR_DIVERGED (-6)  /* solver breakdown / NaN / max_it reached            */

/* precision of a device vector */
#define DGH_F64 0
#define DGH_F32 1

/* Copies the message of the last error on this thread into buf (NUL-terminated). */
int dgh_last_error(char* buf, size_t len);

/* ================================================================ host: stencil tables */

/* Element-type map (childNum, hanging mask) -> eType in [0,18).
 * Replaces GET_ETYPE_BLOCK / getElemType<c> (include/oda/odaUtils.h:334-374,
 * include/oda/odaUtils.txx:13-94, include/oda/cnumEtypes.h:15-229).
 * Returns the eType, or DGH_ERR_MESH for a mask that is not one of the 18 for that child. */
int dgh_etype(int child_num, int hn_mask);

/* Generates the Type2 Laplacian and mass element stencils [8][18][8][8], the Type2
 * restriction stencils [8][18][8][8] and the Type1 restriction stencils [8][8][18][8][8]
 * (coarse child, fine child, coarse eType, coarse vertex, fine vertex) in the reference's
 * index order, from the closed-form hanging-node shape functions.  Replaces the readers
 * createLmatType2 / createMmatType2 (examples/src/backend/handleType2Stencils.C:678-987),
 * readRmatType2Stencil (src/omg/RmatType2StencilsAux.C:15-43) and readRmatType1Stencil
 * (src/omg/RmatType1StencilsAux.C:13-45), which parse the Stencils/ *.inp tables.
 * Any output pointer may be NULL. */
int dgh_stencils_generate(double* lap2, double* mass2, double* rmat2, double* rmat1);

/* ================================================================ host: synthetic points */
#define DGH_PTS_GAUSS 0      /* per-axis N(0.5, 0.1)                                   */
#define DGH_PTS_LOGNORMAL 1  /* per-axis lognormal(0,1) rescaled to [0,1] by min/max   */
#define DGH_PTS_MIXED 2      /* n/2 uniform + n/2 in 8 Gaussian clusters (sd 0.02)     */
#define DGH_PTS_CLUSTERS 3   /* n in 8 Gaussian clusters (sd 0.02)                     */
#define DGH_PTS_UNIFORM 4    /* U(0,1)^3                                               */
/* Writes n points (xyz interleaved) from a xoshiro256** stream seeded by splitmix64(seed);
 * Gaussian variates by Box-Muller.  cluster_seed seeds the cluster centres (MIXED, CLUSTERS).
 * Plays the role of scripts/genGaussSeq.C and scripts/genLogNormSeq.C. */
int dgh_points_generate(int kind, uint64_t n, uint64_t seed, uint64_t cluster_seed, double* xyz);

/* ================================================================ host: octrees */
typedef struct dgh_octree_s* dgh_octree;

/* Points -> complete linear octree (at most max_pts points per leaf, depth <= max_depth) ->
 * 2:1 balanced octree (corner neighbours included when inc_corner != 0).  Applies the
 * driver's point preprocessing first: quantise to max_depth, drop coordinates <= 0 or
 * outside the unit cube, remove duplicate cells (examples/src/drivers/tstMg.C:148-192).
 * Replaces ot::points2Octree (src/oct/Construct.C:727) + ot::balanceOctree
 * (src/oct/Balance.C:31). */
int dgh_octree_from_points(const double* xyz, uint64_t n, uint32_t max_depth, uint32_t max_pts,
                           int inc_corner, dgh_octree* out);
/* Wraps a given sorted, linear, complete octree (anchors xyz[n][3], levels lev[n]). */
int dgh_octree_from_leaves(const uint32_t* xyz, const uint8_t* lev, uint64_t n,
                           uint32_t max_depth, dgh_octree* out);
/* Minimal 2:1 balanced refinement.  Replaces ot::balanceOctree (src/oct/Balance.C:31). */
int dgh_octree_balance(dgh_octree in, int inc_corner, dgh_octree* out);
/* One level of coarsening of complete sibling families.
 * Replaces ot::coarsenOctree (src/oct/Coarsen.C:53-106). */
int dgh_octree_coarsen(dgh_octree in, dgh_octree* out);
int dgh_octree_size(dgh_octree t, uint64_t* n, uint32_t* max_depth);
int dgh_octree_get(dgh_octree t, uint32_t* xyz, uint8_t* lev);
int dgh_octree_free(dgh_octree t);

/* ================================================================ host: DA hierarchy */
typedef struct dgh_damg_s* dgh_damg;

typedef struct {
  uint32_t n_elem;      /* elements (ot::DA::getElementSize)                           */
  uint32_t n_node;      /* owned nodes incl. positive-boundary nodes (getNodeSize +    */
                        /* getBoundaryNodeSize, src/oda/oda.C:530-533)                 */
  uint32_t n_oct;       /* elements + positive-boundary pseudo-octants                 */
  uint32_t max_depth;   /* DA max depth = octree max depth + 1 (odaFactory.C:160)      */
  uint32_t n_hanging;   /* elements with a nonzero hanging mask                        */
} dgh_level_info;

/* Builds the multigrid hierarchy of DAs from the finest balanced octree: coarsen + balance
 * until *nlevels levels exist or a level has fewer than min_coarse octants (the reference
 * uses 500), then the per-level DA (+1 embedding, positive-boundary pseudo-octants, regular
 * node flags, 8-node lists with hanging vertices resolved to the parent's vertex node,
 * Morton child numbers, hanging masks, levels) and the fine-node -> coarse-element transfer
 * maps.  *nlevels returns the actual count.  Replaces ot::DAMGCreateAndSetDA
 * (src/omg/omg.C:1306-2171), the DA constructor (src/oda/oda.C:175-196,
 * src/oda/private/odaFactory.C:100-330, src/oda/private/odaBuildNlist.C:47-1900) and
 * createInterpolationType1 / dummyRestrictMatVecType1 (src/omg/omg.C:2234-2290,
 * src/omg/RestrictMatVec.C:410-484). */
int dgh_damg_build(dgh_octree finest, int* nlevels, uint32_t min_coarse, dgh_damg* out);
int dgh_damg_num_levels(dgh_damg h, int* nlevels);
int dgh_damg_level_info(dgh_damg h, int level, dgh_level_info* info);
/* Element arrays of one level (any pointer may be NULL):
 *   nlist[n_elem][8]  node index of each vertex (the reference's uncompressed m_uiNlist,
 *                     renumbered to the compact node vector, src/oda/oda.C:621-646);
 *   child[n_elem]     Morton child number (include/oda/oda.txx:1087-1109);
 *   hn_mask[n_elem]   hanging-vertex mask (src/oda/oda.C:784-789);
 *   lev[n_elem]       DA level (include/oda/oda.txx:51-56);
 *   anchor[n_elem][3] DA anchor coordinates. */
int dgh_damg_level_arrays(dgh_damg h, int level, uint32_t* nlist, uint8_t* child,
                          uint8_t* hn_mask, uint8_t* lev, uint32_t* anchor);
/* All octants of one level (elements then pseudo-octants, Morton order):
 * anchor[n_oct][3], lev[n_oct], flags[n_oct] (bit 7 = regular node, bit 6 = pseudo-octant). */
int dgh_damg_level_octants(dgh_damg h, int level, uint32_t* anchor, uint8_t* lev,
                           uint8_t* flags);
/* Boundary flags of the nodes of one level: bdy[n_node], 1 for a node on the boundary of the
 * unit cube, else 0.  Replaces assignBoundaryFlags (src/oda/odaUtils.C:1054-1283), whose
 * face/edge/corner codes the Dirichlet operators only test for nonzero
 * (examples/src/backend/omgJac.C:508-586, 859-1160). */
int dgh_damg_level_boundary(dgh_damg h, int level, uint8_t* bdy);
/* Transfer map from level fine_level-1 (coarse) to fine_level: for every fine node, the
 * index of a coarse element whose closure holds it and the node's position code
 * pos = tx + 3 ty + 9 tz (t in {0,1,2}: offset 0, h/2, h inside that coarse element). */
int dgh_damg_transfer_arrays(dgh_damg h, int fine_level, uint32_t* coarse_elem, uint8_t* pos);
/* Per-element coefficients of every level from finest-level values: kappa/mass of a
 * coarse element = same element's value, or the mean of its 8 children taken in child
 * order.  Replaces ASSIGN_MAT_PROP_FINE_TO_COARSE_ELEM_BLOCK
 * (examples/src/backend/matProp.C:192-232).  out_kappa/out_mass hold n_elem of the level. */
int dgh_damg_coarsen_coefficients(dgh_damg h, const double* fine_kappa, const double* fine_mass,
                                  int level, double* out_kappa, double* out_mass);
/* Saves / loads a hierarchy (every level's octants, node numbering, element arrays and the
 * transfer maps) as a versioned binary with a checksum, so that a mesh is built once and
 * loaded by later runs (checkpoint / resume of the mesh stage).  Loading checks the magic,
 * version, byte order, array lengths, index ranges and checksum.  No reference counterpart
 * (the reference rebuilds its DAs from points on every run, src/omg/omg.C:1306). */
int dgh_damg_save(dgh_damg h, const char* path);
int dgh_damg_load(const char* path, dgh_damg* out);
int dgh_damg_free(dgh_damg h);

/* ================================================================ device: runtime */

/* Selects the device and creates the library state (stencil tables, streams).
 * Plays the role of DAMG_Initialize / DA_Initialize (src/omg/omg.C:66, src/oda/odaUtils.C:1742). */
int dgh_init(int device);
int dgh_finalize(void);
/* Device memory and copies for callers without a HIP runtime of their own (the reference
 * keeps its vectors in PETSc Vecs on the host; this is where they cross to the device).
 * No reference counterpart.  Copies are synchronous; dgh_synchronize waits for all work. */
int dgh_alloc(size_t bytes, void** ptr);
int dgh_free(void* ptr);
int dgh_copy_to_device(void* dst_dev, const void* src_host, size_t bytes);
int dgh_copy_to_host(void* dst_host, const void* src_dev, size_t bytes);
int dgh_synchronize(void);

/* ================================================================ device: one mesh level */
typedef struct dgh_mesh_s* dgh_mesh;

/* Uploads one DA level once (host arrays as returned by dgh_damg_level_arrays).  Element
 * data are stored on the device grouped by stencil (Morton order inside a group).
 * Replaces CreateJacobian1 (examples/src/backend/odaJac.C:54-72) and the per-call element
 * decode of Jacobian1MatMult (odaJac.C:290-374). */
int dgh_mesh_create(uint32_t n_elem, uint32_t n_node, uint32_t max_depth, const uint32_t* nlist,
                    const uint8_t* child, const uint8_t* hn_mask, const uint8_t* lev,
                    dgh_mesh* out);
/* Operator coefficients.  kappa == NULL: uniform Laplacian factor lap_fac and mass factor
 * mass_fac (Jacobian1, "-lapMatPropFac"/"-massMatPropFac", odaJac.C:309-313); otherwise the
 * per-element (kappa, mass) of Jacobian2 (examples/src/backend/omgJac.C:1665-1685; mass may
 * be NULL for zero mass). */
int dgh_mesh_set_coefficients(dgh_mesh m, const double* kappa, const double* mass,
                              double lap_fac, double mass_fac);
int dgh_mesh_info(dgh_mesh m, uint32_t* n_elem, uint32_t* n_node);
/* Re-tiles the operator kernel with tiles of at most tile_elems Morton-consecutive elements
 * (default 2048 with 512 threads, or 8192 with 1024 threads on meshes of at least 4M
 * elements; reduced automatically until the tile's LDS fits) and block threads per
 * workgroup (256, 512 or 1024; 0 keeps the current value).  No reference counterpart: the
 * tuning hook of the tile autotuner. */
int dgh_mesh_set_tile(dgh_mesh m, uint32_t tile_elems, uint32_t block);
/* Tiling statistics: tile size, tiles, element records (own + halo), node-table entries,
 * LDS bytes per FP64 workgroup.  No reference counterpart. */
int dgh_mesh_tile_info(dgh_mesh m, uint32_t* tile_elems, uint32_t* n_tiles, uint64_t* n_records,
                       uint64_t* n_node_entries, uint32_t* lds_bytes_f64);
/* Stencil rows of the operator kernels: DGH_ROWS_TYPE = the Laplacian and mass rows read
 * through the scalar cache once per stencil type present in a wavefront; DGH_ROWS_LANE =
 * read per lane from a copy of the stencil table in LDS (one pass per wavefront however many
 * types its lanes hold; pays on small levels, whose small tiles mix many types);
 * DGH_ROWS_CLASS = uniform coefficients only (per-element ones use DGH_ROWS_TYPE): the
 * element matrix lap_fac h/2 L + mass_fac h^3/8 M of each (type, level) class, read through
 * the scalar cache once per class present in a wavefront (half the products per row).
 * Until one is set, each re-tiling picks DGH_ROWS_CLASS for tilings of at least 1024 tiles
 * and DGH_ROWS_TYPE otherwise.  DGH_ROWS_LANE fails if the tiling leaves no LDS room for the
 * table.  No reference counterpart: a kernel choice, also made by dgh_mesh_autotune. */
#define DGH_ROWS_TYPE 0
#define DGH_ROWS_LANE 1
#define DGH_ROWS_CLASS 2
int dgh_mesh_set_lane_stencils(dgh_mesh m, int on);
int dgh_mesh_lane_stencils(dgh_mesh m, int* on);
/* Autotuner: times the operator kernel of precision prec on this mesh for every candidate
 * (tile size, block, stencil rows) choice and keeps the fastest.  Synchronous; uses scratch
 * vectors it allocates itself.  Outputs (any may be NULL): the chosen tile size and block
 * and the time per launch.  No reference counterpart. */
int dgh_mesh_autotune(dgh_mesh m, int prec, uint32_t* tile_elems, uint32_t* block, float* us);
int dgh_mesh_destroy(dgh_mesh m);

/* y = A x.  Replaces Jacobian1MatMult (odaJac.C:290-374) and Jacobian2MatMult
 * (omgJac.C:1687-1775).  prec selects FP64 or FP32 vectors and arithmetic. */
int dgh_matvec(dgh_mesh m, int prec, const void* x_dev, void* y_dev, void* stream);
/* y = A x and xy = x.y (FP64) in one operator launch plus one fixed-order sum of its per-tile
 * partials: the MatMult + VecDot pair of every CG iteration (PETSc KSPSolve_CG behind
 * DAMGSolve, src/omg/omg.C:1166, with -ksp_type cg from bin/options:50-53), for callers that
 * run their own Krylov loop.  xy_dev is one device double.  The partials sit in a buffer the
 * library allocates for the mesh on the first call (and again after a re-tiling that adds
 * tiles). */
int dgh_matvec_dot(dgh_mesh m, const double* x_dev, double* y_dev, double* xy_dev, void* stream);
/* dgh_matvec_dot without any allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight solve loops, harnesses
 * that forbid hipMalloc).  The per-tile partials go to a fixed device array of
 * DGH_NOALLOC_MAX_TILES doubles inside the library (one per GPU, held until the process
 * ends); meshes tiled into more tiles are rejected with DGH_ERR_ARG.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other threads) race. */
#define DGH_NOALLOC_MAX_TILES 16384
int dgh_matvec_dot_noalloc(dgh_mesh m, const double* x_dev, double* y_dev, double* xy_dev, void* stream);
/* d = diag(A).  Replaces Jacobian1MatGetDiagonal (odaJac.C:160-223) and
 * Jacobian2MatGetDiagonal (omgJac.C:1594-1662). */
int dgh_diagonal(dgh_mesh m, int prec, void* d_dev, void* stream);
/* Boundary nodes of the mesh for the Dirichlet operators: bdy[N] host flags (nonzero = on
 * the boundary, as dgh_damg_level_boundary returns them), uploaded once (1 byte per node);
 * NULL removes them.  Replaces SetDirichletJacContexts
 * (examples/src/backend/omgJac.C:508-586). */
int dgh_mesh_set_boundary(dgh_mesh m, const uint8_t* bdy);
/* Dirichlet-masked y = A x: boundary rows are the identity (y_i = x_i) and boundary columns
 * are dropped from the other rows.  Replaces DirichletLaplacianMatMult and
 * DirichletJacobianMatMult (examples/src/backend/omgJac.C:859-1056): coefficients set
 * uniform with lap_fac 1 and mass_fac 0 give the first, mass_fac 1 the second; per-element
 * coefficients apply the same masking to the Jacobian2 operator.  Needs
 * dgh_mesh_set_boundary. */
int dgh_matvec_dirichlet(dgh_mesh m, int prec, const void* x_dev, void* y_dev, void* stream);
/* Diagonal of the Dirichlet-masked operator: 1 on boundary nodes.  Replaces
 * DirichletLaplacianMatGetDiagonal and DirichletJacobianMatGetDiagonal
 * (examples/src/backend/omgJac.C:1058-1160). */
int dgh_diagonal_dirichlet(dgh_mesh m, int prec, void* d_dev, void* stream);
/* Y = A X for k right-hand sides at once (k a multiple of 8, at most 32): X and Y are
 * [N][k] row-major (node-major, right-hand side fastest), 16-byte aligned.  The reference applies
 * Jacobian1MatMult / Jacobian2MatMult (odaJac.C:290-374, omgJac.C:1687-1775) to one vector
 * per call; each column here equals that product.  FP32 runs the element products on the
 * vector ALUs or the matrix cores (the variant of dgh_mesh_set_tile_multi). */
int dgh_matvec_multi(dgh_mesh m, int prec, int k, const void* x_dev, void* y_dev, void* stream);
/* Tile size (elements per tile) of the batched operator, rounded down until its LDS fits,
 * and the FP32 product variant (0: vector ALUs, 8 columns per workgroup slice, 512
 * threads, one column per lane; 1: matrix cores; 2: as 0 with 1024 threads; 3: as 0 with 4
 * columns per slice; 4, 5: as 0 with 2, 4 columns per lane; 6: as 3 with 2 columns per
 * lane).  No reference counterpart (tuning hook). */
int dgh_mesh_set_tile_multi(dgh_mesh m, uint32_t tile_elems, int variant);
/* Times the FP32 batched operator for k columns over the candidate tile sizes and all
 * product variants and keeps the fastest; outputs (may be NULL) the tile size, the variant
 * and its time per launch.  No reference counterpart. */
int dgh_mesh_autotune_multi(dgh_mesh m, int k, uint32_t* tile_elems, int* variant, float* us);

/* ================================================================ device: linear elasticity */
/* Linear elasticity on the mesh, 3 dofs per node (vectors [N][3] interleaved, FP64):
 * y = sum_e fac (mu L_e (x) I3 + (mu + lambda) GD_e) x_e with fac = h/2, L the Laplacian
 * stencils and GD[3k+a][3j+b] = int d_a phi_k d_b phi_j the grad-div stencils; with boundary
 * flags (dgh_mesh_set_boundary) boundary rows return x and boundary columns are dropped.
 * Replaces CreateElasticityMat / ElasticityMatMult (examples/src/backend/elasticityJac.C:129-211,
 * stencils Stencils/GDType2Stencils.inp).  set_elasticity builds the plan once (element order
 * by stencil type, node slot CSR, a [E][24] FP64 buffer of element outputs). */
int dgh_mesh_set_elasticity(dgh_mesh m, double mu, double lambda);
int dgh_matvec_elasticity(dgh_mesh m, const double* x_dev, double* y_dev, void* stream);

/* ================================================================ device: right-hand sides and errors */
/* Functions on the unit cube for right-hand sides, nodal values and errors. */
#define DGH_FN_GAUSSIANS 0  /* three Gaussians exp(-|x - c|^2 / 0.005), c = (0.1,0.1,0.1) on        */
                            /* [0,0.2)^3, (0.3,..) on [0.2,0.4]^3, (0.6,..) on [0.5,0.7]^3        */
                            /* (EVAL_FN, examples/src/backend/omgRhs.C:24-47)                    */
#define DGH_FN_COS_SUM 1    /* sum_{k=1..nfreq} (a0 + 3 a1 k^2 pi^2) cos(k pi x) cos(k pi y)        */
                            /* cos(k pi z): a0 = 1, a1 = 0 is ComputeSol's solution, a0 = a1 = 1  */
                            /* the Helmholtz source of ComputeRHS1 (omgRhs.C:1196-1254, 1340-1430) */
#define DGH_FN_POLY 2       /* c0 + c1 x + c2 y + c3 z + c4 xy + c5 yz + c6 zx + c7 xyz           */
typedef struct {
  int kind;          /* DGH_FN_* */
  int nfreq;         /* COS_SUM: number of frequencies (the reference uses 9) */
  double a0, a1;     /* COS_SUM: mass and Laplacian factors of each term */
  double poly[8];    /* POLY coefficients */
} dgh_function;

/* Element anchors [n_elem][3] (DA coordinates, dgh_damg_level_arrays), uploaded once; needed by
 * the three calls below, which evaluate functions at element points.  The DA's levels and
 * hanging masks come from dgh_mesh_create. */
int dgh_mesh_set_geometry(dgh_mesh m, const uint32_t* anchor);
/* b_i = sum_e (h_e^3 / 8) sum over the 3x3x3 Gauss nodes of w f(x) phi_i(x), phi the element's
 * hanging-node shape functions.  Replaces ComputeRHS2 (examples/src/backend/omgRhs.C:1127-1193)
 * and its shape-function table (Stencils/ShFnStencils.inp), for any dgh_function.  Element
 * contributions are added with device atomics: the last bits may vary from run to run. */
int dgh_rhs_quadrature(dgh_mesh m, const dgh_function* f, double* b_dev, void* stream);
/* u_i = f(node i), from each element's non-hanging vertices.  Replaces ComputeSol
 * (examples/src/backend/omgRhs.C:1196-1254) for any dgh_function. */
int dgh_nodal_values(dgh_mesh m, const dgh_function* f, double* u_dev, void* stream);
/* err = ||f - u_h||_L2 over the mesh by 3x3x3 Gauss quadrature of each element, u_h the finite
 * element function of the nodal values u_dev.  Replaces ComputeError5
 * (examples/src/backend/omgRhs.C:1832-1890, whose exact solution is one fixed cosine) for any
 * dgh_function.  Synchronous; the element sums are reduced in a fixed order. */
int dgh_l2_error(dgh_mesh m, const dgh_function* f, const double* u_dev, double* err);

/* ================================================================ device: intergrid transfer */
typedef struct dgh_transfer_s* dgh_transfer;

/* Uploads the fine-node -> coarse-element map (dgh_damg_transfer_arrays).
 * Replaces createInterpolationType1 (src/omg/omg.C:2234-2290). */
int dgh_transfer_create(dgh_mesh coarse, dgh_mesh fine, const uint32_t* coarse_elem,
                        const uint8_t* pos, dgh_transfer* out);
int dgh_transfer_destroy(dgh_transfer t);
/* c = R f.  Replaces restrictMatVecType1 (src/omg/RestrictMatVec.C:486-606). */
int dgh_restrict(dgh_transfer t, int prec, const void* f_dev, void* c_dev, void* stream);
/* f_out = f_in + R^T c (f_in may equal f_out; f_in == NULL means f_out = R^T c).
 * Replaces prolongMatVecType1 / addProlongMatVec (src/omg/ProlongMatVec.C:32-284). */
int dgh_prolong_add(dgh_transfer t, int prec, const void* c_dev, const void* f_in_dev,
                    void* f_out_dev, void* stream);

/* Jacobian3 product on the coarse level of t: y = A3 x, the coarse operator formed from the
 * fine level's element coefficients.  Replaces Jacobian3MatMult
 * (examples/src/backend/omgJac.C:2139-2339), whose Type2/Type1 element blocks with the fine
 * coefficients sum to the one-level Galerkin product P^T A_f P of the fine mesh's operator (the
 * coarse space is nested in the fine one): it is applied here as prolongation, the fine
 * operator with the fine mesh's coefficients, and restriction.  Two fine-level scratch vectors
 * are allocated for the transfer on the first call of each precision. */
int dgh_matvec_jac3(dgh_transfer t, int prec, const void* x_dev, void* y_dev, void* stream);

/* ================================================================ device: GMG-PCG */
typedef struct dgh_solver_s* dgh_solver;

#define DGH_NORM_UNPRECONDITIONED 0 /* stop on ||r||_2 <= rtol ||b||_2                 */
#define DGH_NORM_PRECONDITIONED 1   /* stop on ||Br||_2 <= rtol ||Br_0||_2 (PETSc CG)  */

#define DGH_SMOOTHER_JACOBI 0
#define DGH_SMOOTHER_LEARNED 1

/* operator of the levels below the finest (-jacType of examples/src/drivers/checkError.C) */
#define DGH_COARSE_JAC2 0   /* each level's own (coarsened) element coefficients           */
#define DGH_COARSE_JAC3 1   /* Jacobian3: the next finer level's element coefficients       */

typedef struct {
  double rtol;          /* relative tolerance (-ksp_rtol)                              */
  double omega;         /* Richardson scale (-mg_levels_ksp_richardson_scale)          */
  int max_it;           /* outer CG iterations (-ksp_max_it)                           */
  int nu_pre;           /* pre-smoothing sweeps (-mg_levels_ksp_max_it)                */
  int nu_post;          /* post-smoothing sweeps                                       */
  int norm_type;        /* DGH_NORM_*                                                  */
  int null_space;       /* 1: constants are the operator's null space (pure Neumann)   */
  int vcycle_prec;      /* DGH_F64: plain FP64; DGH_F32: FP32 V-cycle, FP64 outer CG   */
  int smoother;         /* DGH_SMOOTHER_*                                              */
  int use_graph;        /* capture the V-cycle in a hipGraph and replay it             */
  int coarse_op;        /* DGH_COARSE_*: operator of residuals and sweeps below the finest */
                        /* level and of the assembled coarsest matrix (the Jacobi       */
                        /* diagonals stay each level's own)                             */
  int coarse_dense_max; /* coarsest levels of at most this many nodes (and at most     */
                        /* 40000) are solved with a dense Cholesky inverse; larger ones */
                        /* by Jacobi-preconditioned CG on the device (the reference's   */
                        /* sparse LU, -pc_type lu, is not reproduced)                   */
  double coarse_rtol;   /* iterative coarse solve: stop at ||b - A x|| <= rtol ||b||     */
  int coarse_max_it;    /* iterative coarse solve: iteration cap                        */
} dgh_solver_params;

typedef struct {
  int iterations;       /* CG iterations done                                          */
  int converged;        /* 1 if the stopping test was met                              */
  double rnorm0;        /* initial norm used by the stopping test                      */
  double rnorm;         /* final norm used by the stopping test                        */
  double rel_true;      /* ||b - A x||_2 / ||b||_2 recomputed in FP64 after the solve  */
  double seconds;       /* device time of the solve (hipEvents)                        */
} dgh_solve_stats;

/* Fills p with the reference defaults of bin/options (CG, V(4,4), omega 0.667, rtol 1e-8). */
int dgh_solver_default_params(dgh_solver_params* p);
/* meshes[0..nlevels-1] coarsest first; transfers[i] maps level i to level i+1.  Computes the
 * Jacobi diagonals, assembles the coarsest operator (CreateAndComputeFullJacobian1,
 * odaJac.C:466-536) and factors it on the host.  Replaces DAMGSetKSP + DAMGSetUpLevel
 * (src/omg/omg.C:779-1046, 1171-1303) with the options of bin/options:50-53,182-206. */
int dgh_solver_create(int nlevels, const dgh_mesh* meshes, const dgh_transfer* transfers,
                      const dgh_solver_params* params, dgh_solver* out);
int dgh_solver_destroy(dgh_solver s);
/* Preconditioned CG on the finest level, x_dev = A^{-1} b_dev (x starts at 0).
 * Replaces DAMGSolve -> KSPSolve (src/omg/omg.C:1121-1169).  Synchronous.
 * history (may be NULL) receives up to history_len stopping-test norms, one per iteration
 * starting with the initial one. */
int dgh_solver_solve(dgh_solver s, const double* b_dev, double* x_dev, dgh_solve_stats* stats,
                     double* history, int history_len, void* stream);
/* k independent solves at once (k a multiple of 8, at most 32): B and X are FP64 [N][k]
 * row-major on the device, 16-byte aligned.  Each column runs the KSPSolve CG of dgh_solver_solve with its
 * own scalars and stopping test; all columns share one batched FP32 V-cycle per iteration
 * (FP64 outer CG, flexible update), captured as one hipGraph when params.use_graph is set.
 * A converged column is frozen.  stats: array of k (may be NULL).  Replaces k calls of
 * DAMGSolve (omg.C:1121) on one hierarchy. */
int dgh_solver_solve_multi(dgh_solver s, int k, const double* B_dev, double* X_dev, dgh_solve_stats* stats,
                           void* stream);
/* How the solver solves its coarsest level: *iterative = 0 for the dense inverse, 1 for the
 * Jacobi-preconditioned CG above coarse_dense_max nodes (then *last_iterations = its
 * iterations in the most recent coarse solve).  Either output may be NULL.  No reference
 * counterpart (the reference factors the coarsest matrix with PETSc's sparse LU). */
int dgh_solver_coarse_info(dgh_solver s, int* iterative, int* last_iterations);
/* z = B r: one application of the V-cycle preconditioner (PCApply_MG), FP64 vectors. */
int dgh_solver_vcycle(dgh_solver s, const double* r_dev, double* z_dev, void* stream);

/* ================================================================ learned Jacobi damping
 * No counterpart in the reference (its smoother is damped Jacobi, bin/options:187-193,
 * which is also the fallback here).  A per-element MLP (features: level, hanging type,
 * child number, 8 nodal log-coefficient ratios; 35 -> 64 -> 64 -> 8, ReLU, BF16 weights,
 * FP32 accumulation) predicts one damping logit per element vertex; per node the logits are
 * averaged and omega_n = cap_n * sigmoid(mean), with cap_n = 2 A_nn / sum |element row
 * entries| (every weighted sweep is then an A-norm contraction).  The smoother of a level
 * with learned weights runs x += omega_n / A_nn (b - A x).  Training runs on the device:
 * samples are the errors the solver leaves after pre-smoothing and coarse correction, the
 * objective is their A-norm reduction by nu weighted sweeps, optimiser Adam. */
#define DGH_LEARN_NPARAM 6984
typedef struct {
  int steps;       /* training steps (one level per step, finest first, round robin)   */
  int samples;     /* sample columns per step: a multiple of 8, at most 32              */
  int nu;          /* weighted sweeps in the objective                                  */
  int refresh;     /* steps of a level between regenerations of its samples             */
  float lr;        /* Adam step size                                                    */
  uint64_t seed;   /* sample seeds                                                      */
} dgh_learn_params;
typedef struct {
  int steps;
  double loss_first;   /* mean objective over the levels at their first step           */
  double loss_last;    /* mean objective over the levels at their last step            */
  double seconds;
} dgh_learn_stats;
int dgh_learn_default_params(dgh_learn_params* p);
/* (Re)initialises the network (He-normal hidden weights from seed, output bias at the
 * configured omega). */
int dgh_learned_init(dgh_solver s, uint64_t seed);
/* FP32 master parameters [DGH_LEARN_NPARAM] (host): W1[64][35] b1[64] W2[64][64] b2[64]
 * W3[8][64] b3[8]. */
int dgh_learned_get_params(dgh_solver s, float* theta, int n);
int dgh_learned_set_params(dgh_solver s, const float* theta, int n);
int dgh_learned_train(dgh_solver s, const dgh_learn_params* p, dgh_learn_stats* stats);
/* Per level 1..nlevels-1 (coarse to fine): draws `samples` fresh samples and keeps the
 * learned weights only if their objective beats plain damped Jacobi (params.omega) on them;
 * levels_on (may be NULL) receives the bit mask of levels using learned weights.
 * enable = 0 restores plain Jacobi everywhere.  Solves after this call use the result. */
int dgh_learned_apply(dgh_solver s, int enable, int samples, uint64_t seed, int* levels_on);
/* Test hooks.  eval: network outputs O [E][8] of a level (host, may be NULL) and, when dO
 * [E][8] (host) is given, the gradient of sum(dO . O) [DGH_LEARN_NPARAM] (host, may be
 * NULL).  loss: the objective of a level for weights omega [N] (host; NULL: the network's)
 * on samples [N][S] (host; NULL: drawn from seed), with dL/domega [N] (host, may be NULL). */
int dgh_learned_eval(dgh_solver s, int level, const float* dO, float* O, float* grad);
int dgh_learned_loss(dgh_solver s, int level, const float* omega, const float* samples, int S, int nu,
                     uint64_t seed, double* loss, float* g_omega);

#ifdef __cplusplus
}
#endif

#endif /* DENDRO_HIP_H */
