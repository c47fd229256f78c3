/*
 * dendro_hip.hpp -- C++ facade of libdendrohip with the entry-point names of the reference's
 * omg/oda layer (include/omg/omg.h, examples/include/odaJac.h, examples/include/omgJac.h),
 * header-only over the C ABI of dendro_hip.h.
 *
 * The reference drives one process through a DAMG: an array of per-level contexts
 * (damg[0] coarsest ... damg[nlevels-1] finest) built from the finest balanced octree by
 * DAMGCreateAndSetDA, given a Krylov solver and a right-hand side by DAMGSetKSP, and solved
 * by DAMGSolve.  The same sequence works here; the differences are:
 *   - MPI_Comm arguments are gone (one process, one GPU);
 *   - the operator is fixed (Jacobian1 with lap/mass factors, or Jacobian2 with per-element
 *     coefficients set by DAMGSetCoefficients, the role of SetUserContexts in matProp.C);
 *     DAMGSetKSP therefore takes solver parameters (the options file values) instead of the
 *     matrix-creation callbacks, and a right-hand-side callback that fills a host array;
 *   - vectors (DAMGGetx, DAMGGetRHS) live on the device;
 *   - every function returns the library's status (0, or a negative DGH_ERR_* code with the
 *     message in dgh_last_error), as the reference returns PetscErrorCode.
 */
#ifndef DENDRO_HIP_HPP
#define DENDRO_HIP_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "dendro_hip.h"

namespace dgh {
namespace omg {

/* ot::TreeNode (include/oct/TreeNode.h): anchor and level of one octant. */
struct TreeNode {
  uint32_t x, y, z;
  uint8_t level;
};

struct Hierarchy;

/* One level of the hierarchy (the reference's struct _p_DAMG, include/omg/omg.h). */
struct DAMGLevel {
  int nlevels;              /* number of levels ((ctx)[0]->nlevels)                  */
  int level;                /* 0 = coarsest                                          */
  dgh_level_info info;      /* element / node counts of the level's DA               */
  dgh_mesh da;              /* the level's device operator (its "J")                 */
  double* x;                /* finest level: solution (device, N doubles)            */
  double* b;                /* finest level: right-hand side (device)                */
  void* user;               /* user context, as in the reference                     */
  Hierarchy* h;             /* shared state of all levels                            */
};
typedef DAMGLevel* DAMG;

struct Hierarchy {
  dgh_damg host = nullptr;
  std::vector<dgh_transfer> trans;
  std::vector<dgh_mesh> meshes;
  dgh_solver solver = nullptr;
  dgh_solver_params params{};
  dgh_solve_stats last{};
};

inline std::string lastError() {
  char buf[1024];
  dgh_last_error(buf, sizeof(buf));
  return std::string(buf);
}

#define DGH_TRY(expr)            \
  do {                           \
    const int rc_ = (expr);      \
    if (rc_ < 0) return rc_;     \
  } while (0)

/* DAMG_Initialize (include/omg/omg.h): selects the device and loads the stencil tables. */
inline int DAMG_Initialize(int device = 0) { return dgh_init(device); }
/* DAMG_Finalize */
inline int DAMG_Finalize() { return dgh_finalize(); }

/* DAMGCreateAndSetDA (include/omg/omg.h, src/omg/omg.C:1306): builds nlevels levels from the
 * finest balanced octree (coarsen + balance until nlevels or fewer than minCoarse octants;
 * nlevels returns the actual count), creates every level's DA on the device, the
 * intergrid transfers and the per-level contexts.  Operator coefficients default to
 * Jacobian1 with lap = mass = 1. */
inline int DAMGCreateAndSetDA(int& nlevels, void* user, DAMG** damg, const std::vector<TreeNode>& finestOctree,
                              uint32_t maxDepth, uint32_t minCoarse = 500) {
  if (!damg || finestOctree.empty() || nlevels < 1) return DGH_ERR_ARG;
  std::vector<uint32_t> xyz(finestOctree.size() * 3);
  std::vector<uint8_t> lev(finestOctree.size());
  for (size_t i = 0; i < finestOctree.size(); ++i) {
    xyz[3 * i] = finestOctree[i].x;
    xyz[3 * i + 1] = finestOctree[i].y;
    xyz[3 * i + 2] = finestOctree[i].z;
    lev[i] = finestOctree[i].level;
  }
  dgh_octree oct = nullptr;
  DGH_TRY(dgh_octree_from_leaves(xyz.data(), lev.data(), finestOctree.size(), maxDepth, &oct));
  Hierarchy* H = new Hierarchy();
  int rc = dgh_damg_build(oct, &nlevels, minCoarse, &H->host);
  dgh_octree_free(oct);
  if (rc < 0) {
    delete H;
    return rc;
  }
  DAMG* d = new DAMG[nlevels]();
  for (int l = 0; l < nlevels; ++l) {
    DAMGLevel* L = new DAMGLevel();
    L->nlevels = nlevels;
    L->level = l;
    L->user = user;
    L->h = H;
    d[l] = L;
    if ((rc = dgh_damg_level_info(H->host, l, &L->info)) < 0) break;
    std::vector<uint32_t> nl(size_t(L->info.n_elem) * 8);
    std::vector<uint8_t> ch(L->info.n_elem), hn(L->info.n_elem), lv(L->info.n_elem);
    if ((rc = dgh_damg_level_arrays(H->host, l, nl.data(), ch.data(), hn.data(), lv.data(), nullptr)) < 0) break;
    if ((rc = dgh_mesh_create(L->info.n_elem, L->info.n_node, L->info.max_depth, nl.data(), ch.data(), hn.data(),
                              lv.data(), &L->da)) < 0)
      break;
    H->meshes.push_back(L->da);
    if (l > 0) {
      std::vector<uint32_t> ce(L->info.n_node);
      std::vector<uint8_t> pos(L->info.n_node);
      dgh_transfer t = nullptr;
      if ((rc = dgh_damg_transfer_arrays(H->host, l, ce.data(), pos.data())) < 0) break;
      if ((rc = dgh_transfer_create(d[l - 1]->da, L->da, ce.data(), pos.data(), &t)) < 0) break;
      H->trans.push_back(t);
    }
  }
  DAMGLevel* F = d[nlevels - 1];
  if (rc >= 0) rc = dgh_alloc(sizeof(double) * F->info.n_node, reinterpret_cast<void**>(&F->x));
  if (rc >= 0) rc = dgh_alloc(sizeof(double) * F->info.n_node, reinterpret_cast<void**>(&F->b));
  *damg = d;   // on failure the partial hierarchy is still released by DAMGDestroy
  return rc < 0 ? rc : 0;
}

/* SetUserContexts (examples/src/backend/matProp.C:399-556): per-element coefficients of the
 * finest level, averaged to the coarser levels (Jacobian2).  mass may be NULL. */
inline int DAMGSetCoefficients(DAMG* damg, const double* fineKappa, const double* fineMass) {
  Hierarchy* H = damg[0]->h;
  const int n = damg[0]->nlevels;
  for (int l = 0; l < n; ++l) {
    std::vector<double> k(damg[l]->info.n_elem), m(fineMass ? damg[l]->info.n_elem : 0);
    DGH_TRY(dgh_damg_coarsen_coefficients(H->host, fineKappa, fineMass, l, k.data(), fineMass ? m.data() : nullptr));
    DGH_TRY(dgh_mesh_set_coefficients(damg[l]->da, k.data(), fineMass ? m.data() : nullptr, 1.0, 1.0));
  }
  return 0;
}

/* DAMGSetKSP (include/omg/omg.h, src/omg/omg.C:1171-1303): creates the PCG + V-cycle solver
 * with the given parameters and fills the right-hand side through rhs(damg_finest, b_host). */
inline int DAMGSetKSP(DAMG* damg, const dgh_solver_params& params, int (*rhs)(DAMG, double* b_host)) {
  Hierarchy* H = damg[0]->h;
  const int n = damg[0]->nlevels;
  if (H->solver) dgh_solver_destroy(H->solver);
  H->solver = nullptr;
  H->params = params;
  DGH_TRY(dgh_solver_create(n, H->meshes.data(), H->trans.empty() ? nullptr : H->trans.data(), &params, &H->solver));
  DAMGLevel* F = damg[n - 1];
  if (rhs) {
    std::vector<double> b(F->info.n_node, 0.0);
    DGH_TRY(rhs(F, b.data()));
    DGH_TRY(dgh_copy_to_device(F->b, b.data(), sizeof(double) * b.size()));
  }
  return 0;
}

/* DAMGSolve (src/omg/omg.C:1121): x = A^{-1} b on the finest level, x starting at zero. */
inline int DAMGSolve(DAMG* damg, dgh_solve_stats* stats = nullptr) {
  Hierarchy* H = damg[0]->h;
  DAMGLevel* F = damg[damg[0]->nlevels - 1];
  if (!H->solver) return DGH_ERR_STATE;
  DGH_TRY(dgh_solver_solve(H->solver, F->b, F->x, &H->last, nullptr, 0, nullptr));
  if (stats) *stats = H->last;
  return 0;
}

/* DAMGGetx / DAMGGetRHS (include/omg/omg.h macros): finest-level device vectors. */
inline double* DAMGGetx(DAMG* damg) { return damg[damg[0]->nlevels - 1]->x; }
inline double* DAMGGetRHS(DAMG* damg) { return damg[damg[0]->nlevels - 1]->b; }

/* DAMGDestroy (src/omg/omg.C): frees every level, the solver and the device vectors. */
inline int DAMGDestroy(DAMG* damg) {
  if (!damg || !damg[0]) return 0;
  Hierarchy* H = damg[0]->h;
  const int n = damg[0]->nlevels;
  DAMGLevel* F = damg[n - 1];
  if (F && F->x) dgh_free(F->x);
  if (F && F->b) dgh_free(F->b);
  if (H->solver) dgh_solver_destroy(H->solver);
  for (dgh_transfer t : H->trans) dgh_transfer_destroy(t);
  for (int l = 0; l < n; ++l) {
    if (!damg[l]) continue;
    if (damg[l]->da) dgh_mesh_destroy(damg[l]->da);
    delete damg[l];
  }
  if (H->host) dgh_damg_free(H->host);
  delete H;
  delete[] damg;
  return 0;
}

/* Jacobian1MatMult / Jacobian2MatMult (examples/src/backend/odaJac.C:290, omgJac.C:1775):
 * out = J in on one level, device vectors. */
inline int Jacobian1MatMult(DAMG level, const double* in_dev, double* out_dev) {
  return dgh_matvec(level->da, DGH_F64, in_dev, out_dev, nullptr);
}
inline int Jacobian2MatMult(DAMG level, const double* in_dev, double* out_dev) {
  return dgh_matvec(level->da, DGH_F64, in_dev, out_dev, nullptr);
}
/* Jacobian1MatGetDiagonal / Jacobian2MatGetDiagonal (odaJac.C:160, omgJac.C:1594). */
inline int Jacobian1MatGetDiagonal(DAMG level, double* diag_dev) {
  return dgh_diagonal(level->da, DGH_F64, diag_dev, nullptr);
}
inline int Jacobian2MatGetDiagonal(DAMG level, double* diag_dev) {
  return dgh_diagonal(level->da, DGH_F64, diag_dev, nullptr);
}

#undef DGH_TRY

}  // namespace omg
}  // namespace dgh

#endif
