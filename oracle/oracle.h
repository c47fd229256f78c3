/*
 * oracle.h -- CPU restatement of the reference's matrix-free octree FEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dendro_amd/, include/) may include,
 * link or call this code; it is loaded by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg.  Every function follows the reference
 * source cited beside it, in the reference's loop order and arithmetic.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_stencils         Stencils/ *.inp tables of the reference (tests/golden fixture),
 *                       symmetry / row-sum / partition-of-unity invariants
 *   or_etype            include/oda/cnumEtypes.h (tests/golden fixture), exhaustive 8x256
 *   or_matvec/diagonal  regular-grid closed form (textbook trilinear Q1 element), brute
 *                       force conforming assembly on tiny meshes, A*1 = mass*1
 *   or_restrict/prolong R^T = P exactly, R*1 = 1 on fine nodes, coarse-node injection
 *   or_*_dirichlet      the unmasked or_matvec/or_diagonal with the boundary columns zeroed
 *                       and the boundary rows replaced by the identity (a different loop),
 *                       symmetry of the interior block; boundary flags against the node
 *                       anchors' coordinates
 *   or_diagonal         diag(A) of the constrained trilinear FEM assembly
 *   or_mg_vcycle        a dense textbook multiplicative V-cycle (Richardson + Jacobi
 *                       smoothing from zero guesses, exact coarsest solve) on the assembled
 *                       level matrices and interpolation matrices, for several omega /
 *                       sweep counts and with a constant null space; wrong omega or sweep
 *                       counts fail (negative controls)
 *   or_pcg              iterate k equals the A-norm minimiser over K_k(BA, B r_0) of the
 *                       dense V-cycle B (k = 1..5), stopping norms |r_k| / |B r_k|; another
 *                       preconditioner's minimisers fail (negative control).  PETSc's own
 *                       floating-point order (the reference's KSPCG/PCMG/LU, not in the
 *                       reference tree) is pinned only up to rounding.
 */
#ifndef DENDRO_ORACLE_H
#define DENDRO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int or_etype(int child, int mask);
void or_stencils(double* lap2, double* mass2, double* rmat2, double* rmat1);
/* Type1 element stencils [8 coarse child][18][8 fine child][8][8] (Jacobian3) */
void or_stencils_type1(double* lap1, double* mass1);
/* shape functions at the 3-point Gauss nodes [8][18][8 vertex][3][3][3] (ComputeRHS2) */
void or_shfn(double* shfn);

/* y = A x over one DA level; kappa == NULL selects Jacobian1 with lap_fac/mass_fac. */
void or_matvec(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist,
               const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
               const double* kappa, const double* mass, double lap_fac, double mass_fac,
               const double* lap2, const double* mass2, const double* x, double* y);
/* or_matvec on nthreads threads (all-core CPU baseline); rows may sum in another order */
int or_matvec_mt(int nthreads, uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist,
                 const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                 const double* kappa, const double* mass, double lap_fac, double mass_fac,
                 const double* lap2, const double* mass2, const double* x, double* y);
void or_diagonal(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist,
                 const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                 const double* kappa, const double* mass, double lap_fac, double mass_fac,
                 const double* lap2, const double* mass2, double* d);
void or_matvec_dirichlet(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist,
                         const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                         const uint8_t* bdy, const double* kappa, const double* mass, double lap_fac,
                         double mass_fac, const double* lap2, const double* mass2, const double* x,
                         double* y);
void or_diagonal_dirichlet(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist,
                           const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                           const uint8_t* bdy, const double* kappa, const double* mass,
                           double lap_fac, double mass_fac, const double* lap2, const double* mass2,
                           double* d);

/* One DA level as the transfer loops see it. */
typedef struct {
  uint32_t E, N, max_depth;
  const uint32_t* nlist;   /* [E][8] */
  const uint8_t* child;
  const uint8_t* hnmask;
  const uint8_t* lev;
  const uint32_t* anchor;  /* [E][3] */
} or_level;

/* c = R f (restrictMatVecType1) and f = R^T c (prolongMatVecType1). */
int or_restrict(const or_level* coarse, const or_level* fine, const double* rmat2,
                const double* rmat1, const double* f, double* c);
int or_prolong(const or_level* coarse, const or_level* fine, const double* rmat2,
               const double* rmat1, const double* c, double* f);

/* elasticity (3 dofs per node): grad-div stencils [8][18][24][24] and the product of
 * ElasticityMatMult with optional Dirichlet flags (NULL: none), vectors [N][3] */
void or_stencils_graddiv(double* gd2);
void or_matvec_elasticity(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                          const uint8_t* hnmask, const uint8_t* lev, const uint8_t* bdy, double mu, double lambda,
                          const double* lap2, const double* gd2, const double* x, double* y);

/* functions on the unit cube: kind 0 = ComputeRHS2's Gaussians, 1 = cosine sum
 * sum_{k<=nfreq} (a0 + 3 a1 k^2 pi^2) cos cos cos, 2 = trilinear polynomial poly[8] */
double or_eval_fn(int kind, int nfreq, double a0, double a1, const double* poly, double x, double y, double z);
void or_rhs_quadrature(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                       const uint8_t* hnmask, const uint8_t* lev, const uint32_t* anchor, int kind, int nfreq,
                       double a0, double a1, const double* poly, const double* shfn, double* b);
void or_nodal_values(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* hnmask,
                     const uint8_t* lev, const uint32_t* anchor, int kind, int nfreq, double a0, double a1,
                     const double* poly, double* u);
double or_l2_error(uint32_t E, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                   const uint8_t* hnmask, const uint8_t* lev, const uint32_t* anchor, int kind, int nfreq,
                   double a0, double a1, const double* poly, const double* shfn, const double* u);

/* Jacobian3MatMult (omgJac.C:2139-2233): coarse-level product with the fine level's element
 * coefficients kf/mf (mf NULL: zero mass); returns 0, or -2 when the meshes do not nest. */
int or_matvec_jac3(const or_level* coarse, const or_level* fine, const double* kf, const double* mf,
                   double tol, const double* lap2, const double* mass2, const double* lap1,
                   const double* mass1, const double* x, double* y);

/* Multigrid hierarchy (coarsest level 0) for the V-cycle and PCG. */
typedef struct or_mg_s or_mg;
or_mg* or_mg_create(int nlevels);
void or_mg_free(or_mg* mg);
/* the arrays must stay alive while the hierarchy is used */
int or_mg_set_level(or_mg* mg, int level, uint32_t E, uint32_t N, uint32_t max_depth,
                    const uint32_t* nlist, const uint8_t* child, const uint8_t* hnmask,
                    const uint8_t* lev, const uint32_t* anchor, const double* kappa,
                    const double* mass, double lap_fac, double mass_fac);
int or_mg_setup(or_mg* mg, double omega, int nu_pre, int nu_post, int null_space);
/* factor = 0: assemble the coarsest matrix but leave it unfactored; or_mg_coarse_matrix
 * copies it out (n = or_mg_coarse_size) and or_mg_coarse_set_lu installs an LU factor with
 * LAPACK getrf conventions (row-major L\U, unit L, piv[k] = row swapped with k at step k),
 * e.g. from a blocked library factorisation of a large coarsest level. */
int or_mg_setup_ex(or_mg* mg, double omega, int nu_pre, int nu_post, int null_space, int factor);
int or_mg_coarse_size(or_mg* mg);
/* Jacobian3 (-jacType 3 of examples/src/drivers/checkError.C:658-667) on the levels below the
 * finest, tolerance -tolToCoarsenMatProp; before or_mg_setup */
int or_mg_set_jac3(or_mg* mg, int on, double tol);
/* Dirichlet problem (SetDirichletJacContexts, examples/src/backend/omgJac.C:508-586): the
 * level's operator and diagonal become the Dirichlet-masked ones, transfers drop every entry
 * that touches a boundary node (the suppressed DOFs), the coarsest matrix has identity rows
 * and columns on the boundary; bdy[N] must stay alive; before or_mg_setup */
int or_mg_set_boundary(or_mg* mg, int level, const uint8_t* bdy);
/* level products on nthreads threads (or_matvec_mt); 1 (default) keeps or_matvec */
int or_mg_set_threads(or_mg* mg, int nthreads);
int or_mg_coarse_matrix(or_mg* mg, double* A);
int or_mg_coarse_set_lu(or_mg* mg, const double* lu, const int* piv);
int or_mg_vcycle(or_mg* mg, const double* r, double* z);
/* PETSc-KSPCG restatement; returns iterations, *converged 0/1, history[0..its] norms */
int or_pcg(or_mg* mg, const double* b, double* x, double rtol, int max_it, int norm_type,
           double* history, int history_len, int* converged);
/* y = A x on level `level` of the hierarchy (the level's own operator). */
int or_mg_apply(or_mg* mg, int level, const double* x, double* y);

/* ---- learned Jacobi damping (learned.c; no reference counterpart) ---- */
#define OR_NF 35                 /* element features */
#define OR_NH 64                 /* hidden width (two layers) */
#define OR_NO 8                  /* outputs: one damping logit per element vertex */
#define OR_NPARAM (OR_NH * OR_NF + OR_NH + OR_NH * OR_NH + OR_NH + OR_NO * OR_NH + OR_NO)

/* lkbar[n] = log10(mean kappa over the element vertices at node n), lke[e] = log10(kappa_e) */
void or_learn_node_logk(uint32_t E, uint32_t N, const uint32_t* nlist, const double* kappa, double lap_fac,
                        float* lkbar, float* lke);
/* X[e] = [ (lev-1)/(maxD-1), onehot18(etype), onehot8(child), lkbar[n_k] - lke[e] (k<8) ] */
void or_learn_features(uint32_t E, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                       const uint8_t* hnmask, const uint8_t* lev, const float* lkbar, const float* lke,
                       float* X);
void or_mlp_forward(const float* theta, int round_bf16, uint32_t E, const float* X, double* O);
void or_mlp_backward(const float* theta, int round_bf16, uint32_t E, const float* X, const double* dO,
                     double* grad);
void or_learn_cap(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                  const uint8_t* hnmask, const uint8_t* lev, const double* kappa, const double* mass,
                  double lap_fac, double mass_fac, const double* lap2, const double* mass2, double* cap);
void or_learn_nodal(uint32_t E, uint32_t N, const uint32_t* nlist, const double* O, const double* cap,
                    double* theta_n, double* omega);
void or_learn_nodal_grad(uint32_t E, uint32_t N, const uint32_t* nlist, const double* theta_n, const double* cap,
                         const double* g_omega, double* dO);
double or_learn_loss(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                     const uint8_t* hnmask, const uint8_t* lev, const double* kappa, const double* mass,
                     double lap_fac, double mass_fac, const double* lap2, const double* mass2,
                     const double* omega, const double* invd, int nu, int S, const double* E0, double* g_omega);

#ifdef __cplusplus
}
#endif
#endif
