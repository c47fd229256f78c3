/*
 * oracle.c -- CPU restatement of the reference's octree FEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product.
 *
 * Paths cite the reference tree.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ element types */
/* include/oda/cnumEtypes.h:15-229 (= Matlab/cNumBasedHangingMasks.m), eType order */
static const unsigned char kMask[8][18] = {
    {0, 4, 2, 6, 16, 20, 18, 22, 14, 30, 84, 86, 94, 50, 54, 62, 118, 126},
    {0, 1, 8, 9, 32, 33, 40, 41, 13, 45, 49, 57, 61, 168, 169, 173, 185, 189},
    {0, 8, 1, 9, 64, 72, 65, 73, 11, 75, 200, 201, 203, 81, 89, 91, 217, 219},
    {0, 2, 4, 6, 128, 130, 132, 134, 7, 135, 162, 166, 167, 196, 198, 199, 230, 231},
    {0, 1, 32, 33, 64, 65, 96, 97, 35, 99, 69, 101, 103, 224, 225, 227, 229, 231},
    {0, 2, 128, 130, 16, 18, 144, 146, 138, 154, 19, 147, 155, 208, 210, 218, 211, 219},
    {0, 4, 16, 20, 128, 132, 144, 148, 21, 149, 140, 156, 157, 176, 180, 181, 188, 189},
    {0, 8, 64, 72, 32, 40, 96, 104, 76, 108, 42, 106, 110, 112, 120, 124, 122, 126}};

/* getElemType<cNum> (include/oda/odaUtils.txx:13-94): switch over the 18 masks */
int or_etype(int child, int mask) {
  if (child < 0 || child > 7) return -1;
  for (int t = 0; t < 18; ++t)
    if (kMask[child][t] == mask) return t;
  return -1;
}

/* ------------------------------------------------------------------ stencils */
/* Matlab/parentOfCoarse3d.m: parent vertices of a child's master element */
static const int kEpnuga[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1},
    {1, 1, 1},    {0, -1, -1}, {0, 1, -1},  {-1, 0, -1}, {1, 0, -1}, {0, -1, 1},  {0, 1, 1},
    {-1, 0, 1},   {1, 0, 1},   {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0}, {1, 1, 0},   {0, 0, -1},
    {0, 0, 1},    {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},  {0, 0, 0}};
static const int kMtIdx[8][8] = {{0, 8, 10, 20, 16, 24, 22, 26},  {8, 1, 20, 11, 24, 17, 26, 23},
                                 {10, 20, 2, 9, 22, 26, 18, 25},  {20, 11, 9, 3, 26, 23, 25, 19},
                                 {16, 24, 22, 26, 4, 12, 14, 21}, {24, 17, 26, 23, 12, 5, 21, 15},
                                 {22, 26, 18, 25, 14, 21, 6, 13}, {26, 23, 25, 19, 21, 15, 13, 7}};
static const int kCube[8][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
                                {-1, -1, 1},  {1, -1, 1},  {-1, 1, 1},  {1, 1, 1}};

static void parent_coords(int c, double P[8][3]) {
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) P[i][d] = 4.0 * kEpnuga[kMtIdx[c][i]][d];
  double off[3];
  for (int d = 0; d < 3; ++d) off[d] = P[c][d] - kCube[c][d];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) P[i][d] -= off[d];
}

/* Matlab/coarseShapeFunction3DCoeffs.m: phiCoeffs(i,:) = inv(M) e_i */
static void coarse_coeffs(int c, int mask, double phi[8][8]) {
  double P[8][3], M[8][8], A[8][16];
  parent_coords(c, P);
  for (int i = 0; i < 8; ++i) {
    double X[3];
    for (int d = 0; d < 3; ++d) X[d] = (mask & (1 << i)) ? P[i][d] : kCube[i][d];
    M[i][0] = 1.0;
    M[i][1] = X[0];
    M[i][2] = X[1];
    M[i][3] = X[2];
    M[i][4] = X[0] * X[1];
    M[i][5] = X[1] * X[2];
    M[i][6] = X[2] * X[0];
    M[i][7] = X[0] * X[1] * X[2];
  }
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 16; ++j) A[i][j] = (j < 8) ? M[i][j] : (j - 8 == i ? 1.0 : 0.0);
  for (int col = 0; col < 8; ++col) { /* Gauss-Jordan with partial pivoting */
    int p = col;
    for (int r = col + 1; r < 8; ++r)
      if (fabs(A[r][col]) > fabs(A[p][col])) p = r;
    for (int j = 0; j < 16; ++j) {
      double t = A[col][j];
      A[col][j] = A[p][j];
      A[p][j] = t;
    }
    double iv = 1.0 / A[col][col];
    for (int j = 0; j < 16; ++j) A[col][j] *= iv;
    for (int r = 0; r < 8; ++r)
      if (r != col) {
        double f = A[r][col];
        for (int j = 0; j < 16; ++j) A[r][j] -= f * A[col][j];
      }
  }
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) phi[i][j] = A[j][8 + i]; /* column i of inv(M) */
}

/* Matlab/evalPhi.m, Matlab/evalGradPhi.m */
static double eval_phi(const double* c, double x, double y, double z) {
  return c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * z * x +
         c[7] * x * y * z;
}
static double eval_grad(const double* c, int d, double x, double y, double z) {
  if (d == 0) return c[1] + c[4] * y + c[6] * z + c[7] * y * z;
  if (d == 1) return c[2] + c[4] * x + c[5] * z + c[7] * x * z;
  return c[3] + c[5] * y + c[6] * x + c[7] * x * y;
}

/* Matlab/computeLapStencilsType2.m, computeMassStencilsType2.m (2-point Gauss, weights 1),
 * computeRPstencilsType2.m, computeRPstencilsType1.m */
void or_stencils(double* lap2, double* mass2, double* rmat2, double* rmat1) {
  const double g[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
  static const double xlb[8] = {-1, 0, -1, 0, -1, 0, -1, 0}, xub[8] = {0, 1, 0, 1, 0, 1, 0, 1};
  static const double ylb[8] = {-1, -1, 0, 0, -1, -1, 0, 0}, yub[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  static const double zlb[8] = {-1, -1, -1, -1, 0, 0, 0, 0}, zub[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double phi[8][8];
      coarse_coeffs(c, kMask[c][t], phi);
      size_t base = ((size_t)c * 18 + t) * 64;
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
          double L = 0.0, M = 0.0;
          for (int m = 0; m < 2; ++m)
            for (int n = 0; n < 2; ++n)
              for (int p = 0; p < 2; ++p) {
                for (int d = 0; d < 3; ++d)
                  L += eval_grad(phi[i], d, g[m], g[n], g[p]) * eval_grad(phi[j], d, g[m], g[n], g[p]);
                M += eval_phi(phi[i], g[m], g[n], g[p]) * eval_phi(phi[j], g[m], g[n], g[p]);
              }
          if (lap2) lap2[base + i * 8 + j] = L;
          if (mass2) mass2[base + i * 8 + j] = M;
        }
      if (rmat2)
        for (int i = 0; i < 8; ++i)
          for (int f = 0; f < 8; ++f)
            rmat2[base + i * 8 + f] = eval_phi(phi[i], kCube[f][0], kCube[f][1], kCube[f][2]);
      if (rmat1)
        for (int cf = 0; cf < 8; ++cf) {
          size_t b1 = (((size_t)c * 8 + cf) * 18 + t) * 64;
          double lx = xlb[cf], ux = xub[cf], ly = ylb[cf], uy = yub[cf], lz = zlb[cf], uz = zub[cf];
          double pts[8][3] = {{lx, ly, lz}, {ux, ly, lz}, {lx, uy, lz}, {ux, uy, lz},
                              {lx, ly, uz}, {ux, ly, uz}, {lx, uy, uz}, {ux, uy, uz}};
          for (int i = 0; i < 8; ++i)
            for (int f = 0; f < 8; ++f)
              rmat1[b1 + i * 8 + f] = eval_phi(phi[i], pts[f][0], pts[f][1], pts[f][2]);
        }
    }
}

/* Matlab/computeLapStencilsType1.m, computeMassStencilsType1.m: the coarse element's shape
 * functions integrated over the octant of fine child cf only (2-point Gauss on that octant,
 * volume factor 1/8), lap1/mass1[c][t][cf][i][j] */
void or_stencils_type1(double* lap1, double* mass1) {
  const double g[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double phi[8][8];
      coarse_coeffs(c, kMask[c][t], phi);
      for (int cf = 0; cf < 8; ++cf) {
        /* octant cf of [-1,1]^3 spans [b-1, b] along an axis with child bit b (b = 0: [-1,0]) */
        double mid[3];
        for (int d = 0; d < 3; ++d) mid[d] = ((cf >> d) & 1) ? 0.5 : -0.5;
        size_t base = (((size_t)c * 18 + t) * 8 + cf) * 64;
        for (int i = 0; i < 8; ++i)
          for (int j = 0; j < 8; ++j) {
            double L = 0.0, M = 0.0;
            for (int m = 0; m < 2; ++m)
              for (int n = 0; n < 2; ++n)
                for (int p = 0; p < 2; ++p) {
                  const double X = mid[0] + 0.5 * g[m], Y = mid[1] + 0.5 * g[n], Z = mid[2] + 0.5 * g[p];
                  for (int d = 0; d < 3; ++d) L += eval_grad(phi[i], d, X, Y, Z) * eval_grad(phi[j], d, X, Y, Z);
                  M += eval_phi(phi[i], X, Y, Z) * eval_phi(phi[j], X, Y, Z);
                }
            if (lap1) lap1[base + i * 8 + j] = L / 8.0;
            if (mass1) mass1[base + i * 8 + j] = M / 8.0;
          }
      }
    }
}

/* Matlab/computeShFnMat.m: shfn[c][t][j][m][n][p] = phi_j at the 3-point Gauss nodes
 * (g[m], g[n], g[p]), g = {0, +sqrt(3/5), -sqrt(3/5)} in the order ComputeRHS2 uses them
 * (examples/src/backend/omgRhs.C:1141-1180) */
void or_shfn(double* shfn) {
  const double g[3] = {0.0, sqrt(3.0 / 5.0), -sqrt(3.0 / 5.0)};
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double phi[8][8];
      coarse_coeffs(c, kMask[c][t], phi);
      for (int j = 0; j < 8; ++j)
        for (int m = 0; m < 3; ++m)
          for (int n = 0; n < 3; ++n)
            for (int p = 0; p < 3; ++p)
              shfn[((((size_t)c * 18 + t) * 8 + j) * 3 + m) * 9 + n * 3 + p] = eval_phi(phi[j], g[m], g[n], g[p]);
    }
}

/* ------------------------------------------------------------------ operator */
/* JACOBIAN_TYPE1_MULT_BLOCK (examples/src/backend/odaJac.C:271-288) and
 * JAC_TYPE2_ELEM_MULT_BLOCK (examples/src/backend/omgJac.C:1665-1685):
 * h = hFac 2^(maxD-lev), hFac = 1/2^(maxD-1); fac1 = kappa h/2, fac2 = mass h^3/8;
 * out[idx[k]] += (fac1 L[k][j] + fac2 M[k][j]) in[idx[j]] in element order. */
void or_matvec(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist,
               const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
               const double* kappa, const double* mass, double lap_fac, double mass_fac,
               const double* lap2, const double* mass2, const double* x, double* y) {
  memset(y, 0, sizeof(double) * N);
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  for (uint32_t e = 0; e < E; ++e) {
    const double h = hFac * (double)(1u << (maxD - lev[e]));
    const double k1 = kappa ? kappa[e] : lap_fac;
    const double k2 = kappa ? (mass ? mass[e] : 0.0) : mass_fac;
    const double fac1 = k1 * h / 2.0;
    const double fac2 = k2 * h * h * h / 8.0;
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    const double* L = lap2 + ((size_t)c * 18 + t) * 64;
    const double* M = mass2 + ((size_t)c * 18 + t) * 64;
    const uint32_t* idx = nlist + (size_t)e * 8;
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 8; ++j)
        y[idx[k]] += ((fac1 * L[k * 8 + j]) + (fac2 * M[k * 8 + j])) * x[idx[j]];
  }
}

/* or_matvec on nthreads threads (the all-core CPU baseline; not a reference restatement of
 * its own).  Thread t runs the element loop over a contiguous range of elements.  Node ids
 * follow the Morton order of the node anchors and an element's vertices lie at or after its
 * anchor, so the rows of a range fall mostly in [first node of the range, first node of the
 * next range): those are added in place, the others are kept in a list that the calling
 * thread adds afterwards in thread order.  The sums of a row can come in another order than
 * in or_matvec (rounding-level differences). */
typedef struct {
  uint32_t e0, e1, n0, n1;
  uint32_t maxD;
  const uint32_t* nlist;
  const uint8_t *child, *hnmask, *lev;
  const double *kappa, *mass, *lap2, *mass2, *x;
  double lap_fac, mass_fac;
  double* y;
  uint32_t* dx;     /* deferred rows */
  double* dv;
  size_t nd, cap;
  int fail;
} mt_part;

static void* mt_worker(void* arg) {
  mt_part* P = (mt_part*)arg;
  const double hFac = 1.0 / ((double)(1u << (P->maxD - 1)));
  for (uint32_t e = P->e0; e < P->e1; ++e) {
    const double h = hFac * (double)(1u << (P->maxD - P->lev[e]));
    const double k1 = P->kappa ? P->kappa[e] : P->lap_fac;
    const double k2 = P->kappa ? (P->mass ? P->mass[e] : 0.0) : P->mass_fac;
    const double fac1 = k1 * h / 2.0;
    const double fac2 = k2 * h * h * h / 8.0;
    const int c = P->child[e];
    int t = 0;
    if (P->hnmask[e]) t = or_etype(c, P->hnmask[e]);
    const double* L = P->lap2 + ((size_t)c * 18 + t) * 64;
    const double* M = P->mass2 + ((size_t)c * 18 + t) * 64;
    const uint32_t* idx = P->nlist + (size_t)e * 8;
    for (int k = 0; k < 8; ++k) {
      double s = 0.0;
      for (int j = 0; j < 8; ++j) s += ((fac1 * L[k * 8 + j]) + (fac2 * M[k * 8 + j])) * P->x[idx[j]];
      const uint32_t n = idx[k];
      if (n >= P->n0 && n < P->n1) {
        P->y[n] += s;
      } else {
        if (P->nd == P->cap) {
          size_t nc = P->cap ? 2 * P->cap : 4096;
          uint32_t* dx = (uint32_t*)realloc(P->dx, nc * sizeof(uint32_t));
          double* dv = (double*)realloc(P->dv, nc * sizeof(double));
          if (!dx || !dv) {
            P->dx = dx ? dx : P->dx;
            P->dv = dv ? dv : P->dv;
            P->fail = 1;
            return NULL;
          }
          P->dx = dx;
          P->dv = dv;
          P->cap = nc;
        }
        P->dx[P->nd] = n;
        P->dv[P->nd++] = s;
      }
    }
  }
  return NULL;
}

int or_matvec_mt(int nthreads, uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist,
                 const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                 const double* kappa, const double* mass, double lap_fac, double mass_fac,
                 const double* lap2, const double* mass2, const double* x, double* y) {
  if (nthreads < 1 || nthreads > 256) return -1;
  if ((uint32_t)nthreads > E) nthreads = E ? (int)E : 1;
  memset(y, 0, sizeof(double) * N);
  mt_part* P = (mt_part*)calloc((size_t)nthreads, sizeof(mt_part));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  if (!P || !th) {
    free(P);
    free(th);
    return -1;
  }
  for (int t = 0; t < nthreads; ++t) {
    mt_part* q = &P[t];
    q->e0 = (uint32_t)((uint64_t)E * t / nthreads);
    q->e1 = (uint32_t)((uint64_t)E * (t + 1) / nthreads);
    q->maxD = maxD;
    q->nlist = nlist;
    q->child = child;
    q->hnmask = hnmask;
    q->lev = lev;
    q->kappa = kappa;
    q->mass = mass;
    q->lap2 = lap2;
    q->mass2 = mass2;
    q->x = x;
    q->lap_fac = lap_fac;
    q->mass_fac = mass_fac;
    q->y = y;
    /* first row of the range: the smallest vertex id of its first element */
    uint32_t m = N;
    if (q->e0 < q->e1)
      for (int k = 0; k < 8; ++k) m = nlist[(size_t)q->e0 * 8 + k] < m ? nlist[(size_t)q->e0 * 8 + k] : m;
    q->n0 = (t == 0) ? 0 : m;
  }
  /* row ranges disjoint and increasing: [n0_t, n0_{t+1}) with n0 made non-decreasing */
  for (int t = 1; t < nthreads; ++t)
    if (P[t].n0 < P[t - 1].n0) P[t].n0 = P[t - 1].n0;
  for (int t = 0; t < nthreads; ++t) P[t].n1 = (t + 1 < nthreads) ? P[t + 1].n0 : N;
  int rc = 0, started = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (pthread_create(&th[t], NULL, mt_worker, &P[t])) {
      rc = -1;
      break;
    }
    ++started;
  }
  for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
  for (int t = 0; t < nthreads && rc == 0; ++t) {
    if (P[t].fail) rc = -1;
    for (size_t q = 0; q < P[t].nd; ++q) y[P[t].dx[q]] += P[t].dv[q];
  }
  for (int t = 0; t < nthreads; ++t) {
    free(P[t].dx);
    free(P[t].dv);
  }
  free(P);
  free(th);
  return rc;
}

/* DirichletLaplacianMatMult / DirichletJacobianMatMult (examples/src/backend/omgJac.C:859-1056):
 * the element loop of or_matvec in which a boundary row k is assigned out[idx[k]] = in[idx[k]]
 * and a non-boundary row adds only the columns j whose node is not on the boundary.  The
 * reference forms fac1 = h/2 and fac2 = h^3/8 (uniform coefficients 1); kappa/mass and
 * lap_fac/mass_fac scale them as in or_matvec. */
void or_matvec_dirichlet(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist,
                         const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                         const uint8_t* bdy, const double* kappa, const double* mass, double lap_fac,
                         double mass_fac, const double* lap2, const double* mass2, const double* x,
                         double* y) {
  memset(y, 0, sizeof(double) * N);
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  for (uint32_t e = 0; e < E; ++e) {
    const double h = hFac * (double)(1u << (maxD - lev[e]));
    const double k1 = kappa ? kappa[e] : lap_fac;
    const double k2 = kappa ? (mass ? mass[e] : 0.0) : mass_fac;
    const double fac1 = k1 * h / 2.0;
    const double fac2 = k2 * h * h * h / 8.0;
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    const double* L = lap2 + ((size_t)c * 18 + t) * 64;
    const double* M = mass2 + ((size_t)c * 18 + t) * 64;
    const uint32_t* idx = nlist + (size_t)e * 8;
    for (int k = 0; k < 8; ++k) {
      if (bdy[idx[k]]) {
        y[idx[k]] = x[idx[k]];
        continue;
      }
      for (int j = 0; j < 8; ++j)
        if (!bdy[idx[j]]) y[idx[k]] += ((fac1 * L[k * 8 + j]) + (fac2 * M[k * 8 + j])) * x[idx[j]];
    }
  }
}

/* DirichletLaplacianMatGetDiagonal / DirichletJacobianMatGetDiagonal (omgJac.C:1058-1160):
 * boundary entries are set to 1, the others accumulate the element diagonals. */
void or_diagonal_dirichlet(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist,
                           const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                           const uint8_t* bdy, const double* kappa, const double* mass,
                           double lap_fac, double mass_fac, const double* lap2, const double* mass2,
                           double* d) {
  memset(d, 0, sizeof(double) * N);
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  for (uint32_t e = 0; e < E; ++e) {
    const double h = hFac * (double)(1u << (maxD - lev[e]));
    const double k1 = kappa ? kappa[e] : lap_fac;
    const double k2 = kappa ? (mass ? mass[e] : 0.0) : mass_fac;
    const double fac1 = k1 * h / 2.0;
    const double fac2 = k2 * h * h * h / 8.0;
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    const double* L = lap2 + ((size_t)c * 18 + t) * 64;
    const double* M = mass2 + ((size_t)c * 18 + t) * 64;
    for (int k = 0; k < 8; ++k) {
      const uint32_t n = nlist[(size_t)e * 8 + k];
      if (bdy[n])
        d[n] = 1.0;
      else
        d[n] += ((fac1 * L[k * 9]) + (fac2 * M[k * 9]));
    }
  }
}

/* Jacobian1MatGetDiagonal (odaJac.C:160-223), JAC_TYPE2_ELEM_DIAG_BLOCK (omgJac.C:1594-1612) */
void or_diagonal(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist,
                 const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                 const double* kappa, const double* mass, double lap_fac, double mass_fac,
                 const double* lap2, const double* mass2, double* d) {
  memset(d, 0, sizeof(double) * N);
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  for (uint32_t e = 0; e < E; ++e) {
    const double h = hFac * (double)(1u << (maxD - lev[e]));
    const double k1 = kappa ? kappa[e] : lap_fac;
    const double k2 = kappa ? (mass ? mass[e] : 0.0) : mass_fac;
    const double fac1 = k1 * h / 2.0;
    const double fac2 = k2 * h * h * h / 8.0;
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    const double* L = lap2 + ((size_t)c * 18 + t) * 64;
    const double* M = mass2 + ((size_t)c * 18 + t) * 64;
    for (int k = 0; k < 8; ++k)
      d[nlist[(size_t)e * 8 + k]] += ((fac1 * L[k * 9]) + (fac2 * M[k * 9]));
  }
}

/* ------------------------------------------------------------------ intergrid transfer */
/* The simultaneous coarse/fine loop of INTERGRID_TRANSFER_LOOP_BLOCK
 * (src/omg/RestrictMatVec.C:82-151): advance the fine element until its anchor matches the
 * coarse element's; same level -> Type2 stencil RmatType2[cNumC][ctype], else the 8
 * children with RmatType1[cNumC][cNumF][ctype].  Non-hanging fine vertices only.  The
 * FineTouchedStatus bits (dummyRestrictMatVecType1, RestrictMatVec.C:153-484) let every
 * (fine node, coarse node) pair contribute exactly once; here the first occurrence in loop
 * order is kept (all occurrences carry the same value, the coarse shape function at the
 * fine node). */
typedef struct {
  uint32_t* cap;   /* per fine node: count of taken coarse nodes */
  uint32_t* ids;   /* per fine node: up to kTake taken coarse nodes */
} taken_t;
#define kTake 64

/* The (fine node, coarse node, weight) pairs the loop keeps, in loop order; pairs of weight
 * zero are dropped (they add 0 to a sum: the same value for finite inputs).  The pairs do
 * not depend on the vector, so the multigrid hierarchy builds them once per level. */
typedef struct {
  size_t n, cap;
  uint32_t *fn, *cn;
  double* w;
} pairs_t;

static void pairs_free(pairs_t* P) {
  free(P->fn);
  free(P->cn);
  free(P->w);
  memset(P, 0, sizeof(*P));
}

static int pairs_push(pairs_t* P, uint32_t fn, uint32_t cn, double w) {
  if (P->n == P->cap) {
    size_t nc = P->cap ? 2 * P->cap : 1 << 16;
    uint32_t* a = (uint32_t*)realloc(P->fn, nc * sizeof(uint32_t));
    if (!a) return -1;
    P->fn = a;
    uint32_t* b = (uint32_t*)realloc(P->cn, nc * sizeof(uint32_t));
    if (!b) return -1;
    P->cn = b;
    double* c = (double*)realloc(P->w, nc * sizeof(double));
    if (!c) return -1;
    P->w = c;
    P->cap = nc;
  }
  P->fn[P->n] = fn;
  P->cn[P->n] = cn;
  P->w[P->n++] = w;
  return 0;
}

static int transfer_pairs(const or_level* C, const or_level* F, const double* rmat2,
                          const double* rmat1, const uint8_t* bdyc, const uint8_t* bdyf, pairs_t* out) {
  taken_t tk;
  memset(out, 0, sizeof(*out));
  tk.cap = (uint32_t*)calloc(F->N, sizeof(uint32_t));
  tk.ids = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)F->N * kTake);
  if (!tk.cap || !tk.ids) {
    free(tk.cap);
    free(tk.ids);
    return -1;
  }
  uint32_t f = 0;
  int rc = 0;
  for (uint32_t c = 0; c < C->E && rc == 0; ++c) {
    const uint32_t* ca = C->anchor + 3 * (size_t)c;
    while (f < F->E && (F->anchor[3 * (size_t)f] != ca[0] || F->anchor[3 * (size_t)f + 1] != ca[1] ||
                        F->anchor[3 * (size_t)f + 2] != ca[2]))
      ++f;
    if (f >= F->E) {
      rc = -2;
      break;
    }
    const int cNumC = C->child[c];
    int ctype = 0;
    if (C->hnmask[c]) ctype = or_etype(cNumC, C->hnmask[c]);
    const uint32_t* cIdx = C->nlist + (size_t)c * 8;
    const int nsub = (F->lev[f] == C->lev[c]) ? 1 : 8;
    for (int cNumF = 0; cNumF < nsub && rc == 0; ++cNumF, ++f) {
      const double* R = (nsub == 1) ? rmat2 + ((size_t)cNumC * 18 + ctype) * 64
                                    : rmat1 + (((size_t)cNumC * 8 + cNumF) * 18 + ctype) * 64;
      const uint8_t fh = F->hnmask[f];
      const uint32_t* fIdx = F->nlist + (size_t)f * 8;
      for (int fC = 0; fC < 8 && rc == 0; ++fC) {
        if (fh & (1 << fC)) continue;
        const uint32_t fn = fIdx[fC];
        for (int cC = 0; cC < 8; ++cC) {
          const uint32_t cn = cIdx[cC];
          uint32_t* ids = tk.ids + (size_t)fn * kTake;
          int seen = 0;
          for (uint32_t q = 0; q < tk.cap[fn]; ++q)
            if (ids[q] == cn) {
              seen = 1;
              break;
            }
          if (seen) continue;
          if (tk.cap[fn] >= kTake) {
            rc = -3;
            break;
          }
          ids[tk.cap[fn]++] = cn;
          const double Rval = R[cC * 8 + fC];
          /* ITLB_SET_VALUE_SUPPRESSED_DOFS (RestrictMatVec.C:75-80, ProlongMatVec.C:76-80):
           * no entry between a boundary node and anything */
          if ((bdyf && bdyf[fn]) || (bdyc && bdyc[cn])) continue;
          if (Rval != 0.0 && pairs_push(out, fn, cn, Rval)) {
            rc = -1;
            break;
          }
        }
      }
    }
  }
  free(tk.cap);
  free(tk.ids);
  if (rc) pairs_free(out);
  return rc;
}

/* restrict: out[cn] += w in[fn]; prolong: out[fn] += w in[cn]; pairs in loop order */
static void transfer_apply(const pairs_t* P, uint32_t n_out, const double* in, double* out,
                           int restrict_dir) {
  memset(out, 0, sizeof(double) * n_out);
  if (restrict_dir)
    for (size_t q = 0; q < P->n; ++q) out[P->cn[q]] += P->w[q] * in[P->fn[q]];
  else
    for (size_t q = 0; q < P->n; ++q) out[P->fn[q]] += P->w[q] * in[P->cn[q]];
}

static int transfer(const or_level* C, const or_level* F, const double* rmat2,
                    const double* rmat1, const double* in, double* out, int restrict_dir) {
  pairs_t P;
  int rc = transfer_pairs(C, F, rmat2, rmat1, NULL, NULL, &P);
  if (rc) return rc;
  transfer_apply(&P, restrict_dir ? C->N : F->N, in, out, restrict_dir);
  pairs_free(&P);
  return 0;
}

int or_restrict(const or_level* C, const or_level* F, const double* rmat2, const double* rmat1,
                const double* f, double* c) {
  return transfer(C, F, rmat2, rmat1, f, c, 1);
}
int or_prolong(const or_level* C, const or_level* F, const double* rmat2, const double* rmat1,
               const double* c, double* f) {
  return transfer(C, F, rmat2, rmat1, c, f, 0);
}

/* ------------------------------------------------------------------ elasticity (3 dofs per node) */
/* Matlab/computeGradDivType2.m: gd2[c][t][3k+a][3j+b] = int_{[-1,1]^3} d_a phi_k d_b phi_j
 * (2-point Gauss), the grad-div stencils of Stencils/GDType2Stencils.inp */
void or_stencils_graddiv(double* gd2) {
  const double g[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double phi[8][8];
      coarse_coeffs(c, kMask[c][t], phi);
      double* G = gd2 + ((size_t)c * 18 + t) * 576;
      for (int q = 0; q < 576; ++q) G[q] = 0.0;
      for (int m = 0; m < 2; ++m)
        for (int n = 0; n < 2; ++n)
          for (int p = 0; p < 2; ++p)
            for (int k = 0; k < 8; ++k)
              for (int a = 0; a < 3; ++a) {
                const double ga = eval_grad(phi[k], a, g[m], g[n], g[p]);
                for (int j = 0; j < 8; ++j)
                  for (int b = 0; b < 3; ++b) G[(3 * k + a) * 24 + 3 * j + b] += ga * eval_grad(phi[j], b, g[m], g[n], g[p]);
              }
    }
}

/* ELASTICITY_ELEM_MULT_BLOCK (examples/src/backend/elasticityJac.C:129-165): fac = h/2; a
 * boundary row (bdy nonzero, when bdy is given) is out[3i+d] = in[3i+d]; other rows add, over
 * non-boundary columns j, mu fac L[k][j] in[3j+d] and (mu+lambda) fac GD[3k+d][3j+d'] in[3j+d'];
 * vectors interleaved [N][3]. */
void or_matvec_elasticity(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist, const uint8_t* child,
                          const uint8_t* hnmask, const uint8_t* lev, const uint8_t* bdy, double mu, double lambda,
                          const double* lap2, const double* gd2, const double* x, double* y) {
  memset(y, 0, sizeof(double) * 3 * (size_t)N);
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  for (uint32_t e = 0; e < E; ++e) {
    const double h = hFac * (double)(1u << (maxD - lev[e]));
    const double fac = h / 2.0;
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    const double* L = lap2 + ((size_t)c * 18 + t) * 64;
    const double* G = gd2 + ((size_t)c * 18 + t) * 576;
    const uint32_t* idx = nlist + (size_t)e * 8;
    for (int k = 0; k < 8; ++k) {
      if (bdy && bdy[idx[k]]) {
        for (int d = 0; d < 3; ++d) y[3 * (size_t)idx[k] + d] = x[3 * (size_t)idx[k] + d];
        continue;
      }
      for (int j = 0; j < 8; ++j) {
        if (bdy && bdy[idx[j]]) continue;
        for (int d = 0; d < 3; ++d)
          y[3 * (size_t)idx[k] + d] += (mu * fac * L[k * 8 + j] * x[3 * (size_t)idx[j] + d]);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            y[3 * (size_t)idx[k] + a] += ((mu + lambda) * fac * (G[(3 * k + a) * 24 + 3 * j + b]) * x[3 * (size_t)idx[j] + b]);
      }
    }
  }
}

/* ------------------------------------------------------------------ functions on the cube */
/* EVAL_FN of examples/src/backend/omgRhs.C:24-47 (kind 0), the cosine sums of ComputeSol /
 * ComputeRHS1 (omgRhs.C:1236-1245, 1382-1392) with factors (a0 + 3 a1 k^2 pi^2) (kind 1), and a
 * trilinear polynomial (kind 2, coefficients of 1, x, y, z, xy, yz, zx, xyz). */
typedef struct {
  int kind, nfreq;
  double a0, a1, c[8];
} or_fn;

static double or_eval(const or_fn* f, double x, double y, double z) {
  double r = 0.0;
  if (f->kind == 0) {
    if ((x >= 0.0) && (y >= 0.0) && (z >= 0.0) && (x < 0.2) && (y < 0.2) && (z < 0.2)) {
      double dx = x - 0.1, dy = y - 0.1, dz = z - 0.1;
      r += exp(-((dx * dx + dy * dy + dz * dz) / 0.005));
    }
    if ((x >= 0.2) && (y >= 0.2) && (z >= 0.2) && (x <= 0.4) && (y <= 0.4) && (z <= 0.4)) {
      double dx = x - 0.3, dy = y - 0.3, dz = z - 0.3;
      r += exp(-((dx * dx + dy * dy + dz * dz) / 0.005));
    }
    if ((x >= 0.5) && (y >= 0.5) && (z >= 0.5) && (x <= 0.7) && (y <= 0.7) && (z <= 0.7)) {
      double dx = x - 0.6, dy = y - 0.6, dz = z - 0.6;
      r += exp(-((dx * dx + dy * dy + dz * dz) / 0.005));
    }
    return r;
  }
  if (f->kind == 1) {
    for (int k = 1; k <= f->nfreq; ++k) {
      double fk = k;
      r += (f->a0 + 3 * f->a1 * fk * fk * M_PI * M_PI) * cos(fk * M_PI * x) * cos(fk * M_PI * y) * cos(fk * M_PI * z);
    }
    return r;
  }
  return f->c[0] + f->c[1] * x + f->c[2] * y + f->c[3] * z + f->c[4] * x * y + f->c[5] * y * z + f->c[6] * z * x +
         f->c[7] * x * y * z;
}

double or_eval_fn(int kind, int nfreq, double a0, double a1, const double* poly, double x, double y, double z) {
  or_fn f = {kind, nfreq, a0, a1, {0}};
  if (poly) memcpy(f.c, poly, sizeof(f.c));
  return or_eval(&f, x, y, z);
}

static const double kGw[3] = {8.0 / 9.0, 5.0 / 9.0, 5.0 / 9.0};

/* ComputeRHS2 (examples/src/backend/omgRhs.C:1127-1193) for the function f: per element,
 * 3x3x3 Gauss quadrature of f against the 8 shape functions (shfn[c][t][j][m][n][p]). */
void or_rhs_quadrature(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist, const uint8_t* child,
                       const uint8_t* hnmask, const uint8_t* lev, const uint32_t* anchor, int kind, int nfreq,
                       double a0, double a1, const double* poly, const double* shfn, double* b) {
  or_fn f = {kind, nfreq, a0, a1, {0}};
  if (poly) memcpy(f.c, poly, sizeof(f.c));
  const double g[3] = {0.0, sqrt((3.0 / 5.0)), -sqrt((3.0 / 5.0))};
  memset(b, 0, sizeof(double) * N);
  const double unit = (double)(1u << (maxD - 1));
  for (uint32_t e = 0; e < E; ++e) {
    const double hx = (double)(1u << (maxD - lev[e])) / unit;
    const double x = anchor[3 * (size_t)e] / unit, y = anchor[3 * (size_t)e + 1] / unit,
                 z = anchor[3 * (size_t)e + 2] / unit;
    const double fac = ((hx * hx * hx) / 8.0);
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    for (int j = 0; j < 8; ++j) {
      double integral = 0.0;
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
          for (int p = 0; p < 3; ++p) {
            const double fv = or_eval(&f, (hx * (1.0 + g[m]) * 0.5) + x, (hx * (1.0 + g[n]) * 0.5) + y,
                                      (hx * (1.0 + g[p]) * 0.5) + z);
            integral += kGw[m] * kGw[n] * kGw[p] * fv * shfn[((((size_t)c * 18 + t) * 8 + j) * 3 + m) * 9 + n * 3 + p];
          }
      b[nlist[(size_t)e * 8 + j]] += fac * integral;
    }
  }
}

/* ComputeSol (omgRhs.C:1196-1254) for the function f: values at the non-hanging vertices. */
void or_nodal_values(uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist, const uint8_t* hnmask,
                     const uint8_t* lev, const uint32_t* anchor, int kind, int nfreq, double a0, double a1,
                     const double* poly, double* u) {
  or_fn f = {kind, nfreq, a0, a1, {0}};
  if (poly) memcpy(f.c, poly, sizeof(f.c));
  memset(u, 0, sizeof(double) * N);
  const double unit = (double)(1u << (maxD - 1));
  for (uint32_t e = 0; e < E; ++e) {
    const double hx = (double)(1u << (maxD - lev[e])) / unit;
    const double x = anchor[3 * (size_t)e] / unit, y = anchor[3 * (size_t)e + 1] / unit,
                 z = anchor[3 * (size_t)e + 2] / unit;
    for (int i = 0; i < 8; ++i)
      if (!(hnmask[e] & (1 << i)))
        u[nlist[(size_t)e * 8 + i]] = or_eval(&f, x + (i & 1) * hx, y + ((i >> 1) & 1) * hx, z + ((i >> 2) & 1) * hx);
  }
}

/* ComputeError5 (omgRhs.C:1832-1890) for the function f: sqrt of the sum over elements of the
 * 3x3x3 Gauss quadrature of (f - u_h)^2. */
double or_l2_error(uint32_t E, uint32_t maxD, const uint32_t* nlist, const uint8_t* child, const uint8_t* hnmask,
                   const uint8_t* lev, const uint32_t* anchor, int kind, int nfreq, double a0, double a1,
                   const double* poly, const double* shfn, const double* u) {
  or_fn f = {kind, nfreq, a0, a1, {0}};
  if (poly) memcpy(f.c, poly, sizeof(f.c));
  const double g[3] = {0.0, sqrt((3.0 / 5.0)), -sqrt((3.0 / 5.0))};
  const double unit = (double)(1u << (maxD - 1));
  double error = 0.0;
  for (uint32_t e = 0; e < E; ++e) {
    const double hx = (double)(1u << (maxD - lev[e])) / unit;
    const double x = anchor[3 * (size_t)e] / unit, y = anchor[3 * (size_t)e + 1] / unit,
                 z = anchor[3 * (size_t)e + 2] / unit;
    const double fac = ((hx * hx * hx) / 8.0);
    const int c = child[e];
    int t = 0;
    if (hnmask[e]) t = or_etype(c, hnmask[e]);
    double integral = 0.0;
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n)
        for (int p = 0; p < 3; ++p) {
          const double fv = or_eval(&f, (hx * (1.0 + g[m]) * 0.5) + x, (hx * (1.0 + g[n]) * 0.5) + y,
                                    (hx * (1.0 + g[p]) * 0.5) + z);
          double uh = 0.0;
          for (int j = 0; j < 8; ++j)
            uh += u[nlist[(size_t)e * 8 + j]] * shfn[((((size_t)c * 18 + t) * 8 + j) * 3 + m) * 9 + n * 3 + p];
          integral += kGw[m] * kGw[n] * kGw[p] * (fv - uh) * (fv - uh);
        }
    error += integral * fac;
  }
  return sqrt(error);
}

/* ------------------------------------------------------------------ Jacobian3 */
/* JAC_TYPE3_MULT_BLOCK (examples/src/backend/omgJac.C:2139-2233): the coarse operator formed
 * from the next finer level's element coefficients.  The coarse and fine element loops run
 * together (the fine element advanced until its anchor matches the coarse one).  A coarse
 * element that is also a fine element takes the Type2 stencils of its (child, eType) with that
 * fine element's coefficients; one with 8 fine children takes the Type2 stencils with their
 * mean coefficients when max - min <= tol max for both, else the sum over the children of the
 * Type1 stencils [child][eType][fine child] with each child's coefficients.  kf/mf: fine
 * element coefficients (mf NULL: zero mass). */
int or_matvec_jac3(const or_level* C, const or_level* F, const double* kf, const double* mf, double tol,
                   const double* lap2, const double* mass2, const double* lap1, const double* mass1,
                   const double* x, double* y) {
  memset(y, 0, sizeof(double) * C->N);
  const uint32_t maxD = C->max_depth;
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  uint32_t f = 0;
  for (uint32_t c = 0; c < C->E; ++c) {
    const uint32_t* ca = C->anchor + 3 * (size_t)c;
    while (f < F->E && (F->anchor[3 * (size_t)f] != ca[0] || F->anchor[3 * (size_t)f + 1] != ca[1] ||
                        F->anchor[3 * (size_t)f + 2] != ca[2]))
      ++f;
    if (f >= F->E) return -2;
    const unsigned lev = C->lev[c];
    const double h = hFac * (double)(1u << (maxD - lev));
    const uint32_t* idx = C->nlist + (size_t)c * 8;
    const int cn = C->child[c];
    int t = 0;
    if (C->hnmask[c]) t = or_etype(cn, C->hnmask[c]);
    const double* L2 = lap2 + ((size_t)cn * 18 + t) * 64;
    const double* M2 = mass2 + ((size_t)cn * 18 + t) * 64;
    if (F->lev[f] == lev) {
      const double fac1 = kf[f] * h / 2.0;
      const double fac2 = (mf ? mf[f] : 0.0) * h * h * h / 8.0;
      for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 8; ++j) y[idx[k]] += ((fac1 * L2[k * 8 + j]) + (fac2 * M2[k * 8 + j])) * x[idx[j]];
      ++f;
      continue;
    }
    if (f + 8 > F->E) return -2;
    double p1[8], p2[8];
    for (int q = 0; q < 8; ++q, ++f) {
      p1[q] = kf[f];
      p2[q] = mf ? mf[f] : 0.0;
    }
    double max1 = p1[0], min1 = p1[0], max2 = p2[0], min2 = p2[0], tot1 = 0.0, tot2 = 0.0;
    for (int q = 0; q < 8; ++q) {
      if (p1[q] > max1) max1 = p1[q];
      if (p2[q] > max2) max2 = p2[q];
      if (p1[q] < min1) min1 = p1[q];
      if (p2[q] < min2) min2 = p2[q];
      tot1 += p1[q];
      tot2 += p2[q];
    }
    if (((max1 - min1) <= (tol * max1)) && ((max2 - min2) <= (tol * max2))) {
      const double fac1 = tot1 * h / 16.0;
      const double fac2 = tot2 * h * h * h / 64.0;
      for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 8; ++j) y[idx[k]] += ((fac1 * L2[k * 8 + j]) + (fac2 * M2[k * 8 + j])) * x[idx[j]];
    } else {
      for (int q = 0; q < 8; ++q) {
        const double fac1 = p1[q] * h / 2.0;
        const double fac2 = p2[q] * h * h * h / 8.0;
        const double* L1 = lap1 + ((((size_t)cn * 18 + t) * 8 + q) * 64);
        const double* M1 = mass1 + ((((size_t)cn * 18 + t) * 8 + q) * 64);
        for (int k = 0; k < 8; ++k)
          for (int j = 0; j < 8; ++j) y[idx[k]] += ((fac1 * L1[k * 8 + j]) + (fac2 * M1[k * 8 + j])) * x[idx[j]];
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ multigrid + CG */
typedef struct {
  or_level lv;
  const double *kappa, *mass;
  double lap_fac, mass_fac;
  double* invdiag;
  const uint8_t* bdy;      /* Dirichlet: boundary flags of the nodes (NULL: natural) */
  double *b, *x, *r, *z;   /* work vectors of this level */
  pairs_t tp;              /* transfer pairs from the next coarser level (level >= 1) */
} mg_level;

struct or_mg_s {
  int nlevels;
  mg_level* L;
  double *lap2, *mass2, *rmat2, *rmat1;
  double omega;
  int nu_pre, nu_post, null_space;
  /* coarse LU (dense, partial pivoting) */
  int nc;
  double* lu;
  int* piv;
  int factored;
  int nthreads;  /* > 1: level products by or_matvec_mt */
  int jac3;      /* Jacobian3 on the levels below the finest (operator of residuals and sweeps,
                    and the assembled coarsest matrix); the Jacobi diagonals stay Jacobian2's */
  double *lap1, *mass1;
  double jac3_tol;
};

or_mg* or_mg_create(int nlevels) {
  if (nlevels < 1) return NULL;
  or_mg* mg = (or_mg*)calloc(1, sizeof(or_mg));
  mg->nlevels = nlevels;
  mg->L = (mg_level*)calloc(nlevels, sizeof(mg_level));
  mg->lap2 = (double*)malloc(sizeof(double) * 9216);
  mg->mass2 = (double*)malloc(sizeof(double) * 9216);
  mg->rmat2 = (double*)malloc(sizeof(double) * 9216);
  mg->rmat1 = (double*)malloc(sizeof(double) * 73728);
  or_stencils(mg->lap2, mg->mass2, mg->rmat2, mg->rmat1);
  mg->lap1 = (double*)malloc(sizeof(double) * 73728);
  mg->mass1 = (double*)malloc(sizeof(double) * 73728);
  or_stencils_type1(mg->lap1, mg->mass1);
  return mg;
}

void or_mg_free(or_mg* mg) {
  if (!mg) return;
  for (int l = 0; l < mg->nlevels; ++l) {
    free(mg->L[l].invdiag);
    free(mg->L[l].b);
    free(mg->L[l].x);
    free(mg->L[l].r);
    free(mg->L[l].z);
    pairs_free(&mg->L[l].tp);
  }
  free(mg->L);
  free(mg->lap2);
  free(mg->mass2);
  free(mg->rmat2);
  free(mg->rmat1);
  free(mg->lap1);
  free(mg->mass1);
  free(mg->lu);
  free(mg->piv);
  free(mg);
}

int or_mg_set_level(or_mg* mg, int l, uint32_t E, uint32_t N, uint32_t maxD, const uint32_t* nlist,
                    const uint8_t* child, const uint8_t* hnmask, const uint8_t* lev,
                    const uint32_t* anchor, const double* kappa, const double* mass, double lap_fac,
                    double mass_fac) {
  if (!mg || l < 0 || l >= mg->nlevels) return -1;
  mg_level* L = &mg->L[l];
  L->lv.E = E;
  L->lv.N = N;
  L->lv.max_depth = maxD;
  L->lv.nlist = nlist;
  L->lv.child = child;
  L->lv.hnmask = hnmask;
  L->lv.lev = lev;
  L->lv.anchor = anchor;
  L->kappa = kappa;
  L->mass = mass;
  L->lap_fac = lap_fac;
  L->mass_fac = mass_fac;
  return 0;
}

static void apply_A(or_mg* mg, int l, const double* x, double* y) {
  mg_level* L = &mg->L[l];
  if (L->bdy) {
    or_matvec_dirichlet(L->lv.E, L->lv.N, L->lv.max_depth, L->lv.nlist, L->lv.child, L->lv.hnmask, L->lv.lev,
                        L->bdy, L->kappa, L->mass, L->lap_fac, L->mass_fac, mg->lap2, mg->mass2, x, y);
    return;
  }
  if (mg->nthreads > 1 &&
      or_matvec_mt(mg->nthreads, L->lv.E, L->lv.N, L->lv.max_depth, L->lv.nlist, L->lv.child,
                   L->lv.hnmask, L->lv.lev, L->kappa, L->mass, L->lap_fac, L->mass_fac, mg->lap2,
                   mg->mass2, x, y) == 0)
    return;
  or_matvec(L->lv.E, L->lv.N, L->lv.max_depth, L->lv.nlist, L->lv.child, L->lv.hnmask, L->lv.lev,
            L->kappa, L->mass, L->lap_fac, L->mass_fac, mg->lap2, mg->mass2, x, y);
}

/* The operator of level l inside the V-cycle (PETSc's Jmat of the level): with Jacobian3 the
 * levels below the finest use the next finer level's element coefficients
 * (examples/src/drivers/checkError.C:658-667 wires them; Jacobian3MatMult, omgJac.C:2235-2339);
 * uniform coefficients make Jacobian3 equal to the level's own operator. */
static int j3_level(const or_mg* mg, int l) {
  return mg->jac3 && l < mg->nlevels - 1 && mg->L[l + 1].kappa != NULL;
}
static void apply_J(or_mg* mg, int l, const double* x, double* y) {
  if (!j3_level(mg, l)) {
    apply_A(mg, l, x, y);
    return;
  }
  or_matvec_jac3(&mg->L[l].lv, &mg->L[l + 1].lv, mg->L[l + 1].kappa, mg->L[l + 1].mass, mg->jac3_tol, mg->lap2,
                 mg->mass2, mg->lap1, mg->mass1, x, y);
}

int or_mg_apply(or_mg* mg, int level, const double* x, double* y) {
  if (!mg || level < 0 || level >= mg->nlevels) return -1;
  apply_A(mg, level, x, y);
  return 0;
}

static void project_mean(double* v, uint32_t n) {
  double s = 0.0;
  for (uint32_t i = 0; i < n; ++i) s += v[i];
  s /= (double)n;
  for (uint32_t i = 0; i < n; ++i) v[i] -= s;
}

/* Coarsest operator assembled from the same element blocks (CreateAndComputeFullJacobian1,
 * odaJac.C:466-536; BUILD_FULL_JAC_TYPE2_BLOCK, omgJac.C:1803-1858) and factored densely
 * (the reference's -mg_coarse_pc_type lu, bin/options:203-206).  With a constant null space
 * the factored matrix is A + a 11^T (a = trace/n^2) applied to mean-free right-hand sides,
 * which returns the mean-free solution of A x = b. */
static int coarse_assemble_jac3(or_mg* mg);
static int coarse_assemble(or_mg* mg) {
  mg_level* L = &mg->L[0];
  const int n = (int)L->lv.N;
  mg->nc = n;
  mg->factored = 0;
  mg->lu = (double*)calloc((size_t)n * n, sizeof(double));
  mg->piv = (int*)malloc(sizeof(int) * n);
  if (!mg->lu || !mg->piv) return -1;
  if (j3_level(mg, 0)) {
    int rc = coarse_assemble_jac3(mg);
    if (rc) return rc;
    goto null_space;
  }
  /* element by element: entry (idx[k], idx[j]) += fac1 L[k][j] + fac2 M[k][j], elements in
   * storage order -- the same terms in the same order as the column products A e_j */
  const double hFac = 1.0 / ((double)(1u << (L->lv.max_depth - 1)));
  for (uint32_t e = 0; e < L->lv.E; ++e) {
    const double h = hFac * (double)(1u << (L->lv.max_depth - L->lv.lev[e]));
    const double k1 = L->kappa ? L->kappa[e] : L->lap_fac;
    const double k2 = L->kappa ? (L->mass ? L->mass[e] : 0.0) : L->mass_fac;
    const double fac1 = k1 * h / 2.0;
    const double fac2 = k2 * h * h * h / 8.0;
    const int c = L->lv.child[e];
    int t = 0;
    if (L->lv.hnmask[e]) t = or_etype(c, L->lv.hnmask[e]);
    const double* Ls = mg->lap2 + ((size_t)c * 18 + t) * 64;
    const double* Ms = mg->mass2 + ((size_t)c * 18 + t) * 64;
    const uint32_t* idx = L->lv.nlist + (size_t)e * 8;
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 8; ++j)
        mg->lu[(size_t)idx[k] * n + idx[j]] += ((fac1 * Ls[k * 8 + j]) + (fac2 * Ms[k * 8 + j])) * 1.0;
  }
null_space:
  if (L->bdy) { /* ComputeDirichletJacobian (omgJac.C:382-506): boundary rows and columns of the
                   identity */
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (L->bdy[i] || L->bdy[j]) mg->lu[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
  }
  if (mg->null_space) {
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += mg->lu[(size_t)i * n + i];
    const double a = tr / ((double)n * n);
    for (size_t q = 0; q < (size_t)n * n; ++q) mg->lu[q] += a;
  }
  return 0;
}

/* BUILD_FULL_JAC_TYPE3_BLOCK (examples/src/backend/omgJac.C:2341-2448): the coarsest matrix
 * from level 1's element coefficients, element by element with the blocks of
 * or_matvec_jac3 (Type2 for a coarse element that is also a fine one or whose 8 children
 * agree within tol, else the sum of its children's Type1 blocks). */
static int coarse_assemble_jac3(or_mg* mg) {
  const or_level* C = &mg->L[0].lv;
  const or_level* F = &mg->L[1].lv;
  const double* kf = mg->L[1].kappa;
  const double* mf = mg->L[1].mass;
  const int n = (int)C->N;
  double* A = mg->lu;
  const uint32_t maxD = C->max_depth;
  const double hFac = 1.0 / ((double)(1u << (maxD - 1)));
  uint32_t f = 0;
  for (uint32_t c = 0; c < C->E; ++c) {
    const uint32_t* ca = C->anchor + 3 * (size_t)c;
    while (f < F->E && (F->anchor[3 * (size_t)f] != ca[0] || F->anchor[3 * (size_t)f + 1] != ca[1] ||
                        F->anchor[3 * (size_t)f + 2] != ca[2]))
      ++f;
    if (f >= F->E) return -2;
    const double h = hFac * (double)(1u << (maxD - C->lev[c]));
    const uint32_t* idx = C->nlist + (size_t)c * 8;
    const int cn = C->child[c];
    int t = 0;
    if (C->hnmask[c]) t = or_etype(cn, C->hnmask[c]);
    const double* L2 = mg->lap2 + ((size_t)cn * 18 + t) * 64;
    const double* M2 = mg->mass2 + ((size_t)cn * 18 + t) * 64;
    double K[64];
    if (F->lev[f] == C->lev[c]) {
      const double fac1 = kf[f] * h / 2.0, fac2 = (mf ? mf[f] : 0.0) * h * h * h / 8.0;
      for (int q = 0; q < 64; ++q) K[q] = (fac1 * L2[q]) + (fac2 * M2[q]);
      ++f;
    } else {
      if (f + 8 > F->E) return -2;
      double p1[8], p2[8], max1, min1, max2, min2, tot1 = 0.0, tot2 = 0.0;
      for (int q = 0; q < 8; ++q, ++f) {
        p1[q] = kf[f];
        p2[q] = mf ? mf[f] : 0.0;
      }
      max1 = min1 = p1[0];
      max2 = min2 = p2[0];
      for (int q = 0; q < 8; ++q) {
        max1 = p1[q] > max1 ? p1[q] : max1;
        min1 = p1[q] < min1 ? p1[q] : min1;
        max2 = p2[q] > max2 ? p2[q] : max2;
        min2 = p2[q] < min2 ? p2[q] : min2;
        tot1 += p1[q];
        tot2 += p2[q];
      }
      if (((max1 - min1) <= (mg->jac3_tol * max1)) && ((max2 - min2) <= (mg->jac3_tol * max2))) {
        const double fac1 = tot1 * h / 16.0, fac2 = tot2 * h * h * h / 64.0;
        for (int q = 0; q < 64; ++q) K[q] = (fac1 * L2[q]) + (fac2 * M2[q]);
      } else {
        for (int q = 0; q < 64; ++q) K[q] = 0.0;
        for (int cf = 0; cf < 8; ++cf) {
          const double fac1 = p1[cf] * h / 2.0, fac2 = p2[cf] * h * h * h / 8.0;
          const double* L1 = mg->lap1 + (((size_t)cn * 18 + t) * 8 + cf) * 64;
          const double* M1 = mg->mass1 + (((size_t)cn * 18 + t) * 8 + cf) * 64;
          for (int q = 0; q < 64; ++q) K[q] += (fac1 * L1[q]) + (fac2 * M1[q]);
        }
      }
    }
    for (int k = 0; k < 8; ++k)
      for (int j = 0; j < 8; ++j) A[(size_t)idx[k] * n + idx[j]] += K[k * 8 + j];
  }
  return 0;
}

static int coarse_factor(or_mg* mg) {
  const int n = mg->nc;
  for (int k = 0; k < n; ++k) { /* LU with partial pivoting, row-major */
    int p = k;
    double best = fabs(mg->lu[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(mg->lu[(size_t)i * n + k]) > best) {
        best = fabs(mg->lu[(size_t)i * n + k]);
        p = i;
      }
    mg->piv[k] = p;
    if (best == 0.0) return -2;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = mg->lu[(size_t)k * n + j];
        mg->lu[(size_t)k * n + j] = mg->lu[(size_t)p * n + j];
        mg->lu[(size_t)p * n + j] = t;
      }
    const double iv = 1.0 / mg->lu[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double* ri = mg->lu + (size_t)i * n;
      const double f = ri[k] * iv;
      ri[k] = f;
      if (f == 0.0) continue;
      const double* rk = mg->lu + (size_t)k * n;
      for (int j = k + 1; j < n; ++j) ri[j] -= f * rk[j];
    }
  }
  mg->factored = 1;
  return 0;
}

static void coarse_solve(or_mg* mg, const double* b, double* x) {
  const int n = mg->nc;
  for (int i = 0; i < n; ++i) x[i] = b[i];
  if (mg->null_space) project_mean(x, (uint32_t)n);
  for (int k = 0; k < n; ++k)
    if (mg->piv[k] != k) {
      double t = x[k];
      x[k] = x[mg->piv[k]];
      x[mg->piv[k]] = t;
    }
  for (int i = 0; i < n; ++i) {
    double s = x[i];
    const double* ri = mg->lu + (size_t)i * n;
    for (int j = 0; j < i; ++j) s -= ri[j] * x[j];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    const double* ri = mg->lu + (size_t)i * n;
    for (int j = i + 1; j < n; ++j) s -= ri[j] * x[j];
    x[i] = s / ri[i];
  }
}

int or_mg_setup(or_mg* mg, double omega, int nu_pre, int nu_post, int null_space) {
  return or_mg_setup_ex(mg, omega, nu_pre, nu_post, null_space, 1);
}

int or_mg_setup_ex(or_mg* mg, double omega, int nu_pre, int nu_post, int null_space, int factor) {
  if (!mg) return -1;
  mg->omega = omega;
  mg->nu_pre = nu_pre;
  mg->nu_post = nu_post;
  mg->null_space = null_space;
  for (int l = 0; l < mg->nlevels; ++l) {
    mg_level* L = &mg->L[l];
    const uint32_t N = L->lv.N;
    L->invdiag = (double*)malloc(sizeof(double) * N);
    L->b = (double*)calloc(N, sizeof(double));
    L->x = (double*)calloc(N, sizeof(double));
    L->r = (double*)calloc(N, sizeof(double));
    L->z = (double*)calloc(N, sizeof(double));
    if (L->bdy)
      or_diagonal_dirichlet(L->lv.E, N, L->lv.max_depth, L->lv.nlist, L->lv.child, L->lv.hnmask, L->lv.lev,
                            L->bdy, L->kappa, L->mass, L->lap_fac, L->mass_fac, mg->lap2, mg->mass2, L->invdiag);
    else
      or_diagonal(L->lv.E, N, L->lv.max_depth, L->lv.nlist, L->lv.child, L->lv.hnmask, L->lv.lev,
                  L->kappa, L->mass, L->lap_fac, L->mass_fac, mg->lap2, mg->mass2, L->invdiag);
    /* PCSetUp_Jacobi: the reciprocal of the diagonal */
    for (uint32_t i = 0; i < N; ++i) L->invdiag[i] = 1.0 / L->invdiag[i];
  }
  for (int l = 1; l < mg->nlevels; ++l) {
    pairs_free(&mg->L[l].tp);
    int rp = transfer_pairs(&mg->L[l - 1].lv, &mg->L[l].lv, mg->rmat2, mg->rmat1, mg->L[l - 1].bdy, mg->L[l].bdy,
                            &mg->L[l].tp);
    if (rp) return rp;
  }
  int rc = coarse_assemble(mg);
  if (rc || !factor) return rc;
  return coarse_factor(mg);
}

int or_mg_set_boundary(or_mg* mg, int level, const uint8_t* bdy) {
  if (!mg || level < 0 || level >= mg->nlevels) return -1;
  mg->L[level].bdy = bdy;
  return 0;
}

int or_mg_set_threads(or_mg* mg, int nthreads) {
  if (!mg || nthreads < 1 || nthreads > 256) return -1;
  mg->nthreads = nthreads;
  return 0;
}

int or_mg_set_jac3(or_mg* mg, int on, double tol) {
  if (!mg || tol < 0.0) return -1;
  mg->jac3 = on;
  mg->jac3_tol = tol;
  return 0;
}

int or_mg_coarse_size(or_mg* mg) { return mg ? mg->nc : -1; }

int or_mg_coarse_matrix(or_mg* mg, double* A) {
  if (!mg || !mg->lu || mg->factored) return -1;
  memcpy(A, mg->lu, sizeof(double) * (size_t)mg->nc * mg->nc);
  return 0;
}

int or_mg_coarse_set_lu(or_mg* mg, const double* lu, const int* piv) {
  if (!mg || !mg->lu) return -1;
  memcpy(mg->lu, lu, sizeof(double) * (size_t)mg->nc * mg->nc);
  memcpy(mg->piv, piv, sizeof(int) * mg->nc);
  mg->factored = 1;
  return 0;
}

/* KSPRICHARDSON + PCJACOBI, norm type none (bin/options:187-193):
 *   r = b (zero guess) or b - A x;  repeat nu: x += omega * invdiag .* r; r = b - A x
 * (the residual after the last sweep is not formed). */
static void richardson(or_mg* mg, int l, const double* b, double* x, int nu, int guess_zero) {
  mg_level* L = &mg->L[l];
  const uint32_t N = L->lv.N;
  double* r = L->r;
  if (guess_zero) {
    memcpy(r, b, sizeof(double) * N);
  } else {
    apply_J(mg, l, x, r);
    for (uint32_t i = 0; i < N; ++i) r[i] = b[i] - r[i];
  }
  for (int it = 0; it < nu; ++it) {
    for (uint32_t i = 0; i < N; ++i) x[i] = x[i] + mg->omega * (r[i] * L->invdiag[i]);
    if (it + 1 < nu) {
      apply_J(mg, l, x, r);
      for (uint32_t i = 0; i < N; ++i) r[i] = b[i] - r[i];
    }
  }
}

/* PCMGMCycle_Private, multiplicative V-cycle: pre-smooth (zero guess), residual, restrict,
 * zero coarse guess, recurse (coarsest: direct solve), interpolate-add, post-smooth. */
static int vcycle(or_mg* mg, int l, const double* b, double* x) {
  if (l == 0) {
    coarse_solve(mg, b, x);
    return 0;
  }
  mg_level* L = &mg->L[l];
  mg_level* Lc = &mg->L[l - 1];
  const uint32_t N = L->lv.N;
  memset(x, 0, sizeof(double) * N);
  richardson(mg, l, b, x, mg->nu_pre, 1);
  apply_J(mg, l, x, L->r);
  for (uint32_t i = 0; i < N; ++i) L->r[i] = b[i] - L->r[i];
  transfer_apply(&L->tp, Lc->lv.N, L->r, Lc->b, 1);
  memset(Lc->x, 0, sizeof(double) * Lc->lv.N);
  int rc = vcycle(mg, l - 1, Lc->b, Lc->x);
  if (rc) return rc;
  transfer_apply(&L->tp, N, Lc->x, L->z, 0);
  for (uint32_t i = 0; i < N; ++i) x[i] = x[i] + L->z[i];
  richardson(mg, l, b, x, mg->nu_post, 0);
  return 0;
}

int or_mg_vcycle(or_mg* mg, const double* r, double* z) {
  if (!mg) return -1;
  if (!mg->factored) return -4;
  const int top = mg->nlevels - 1;
  int rc = vcycle(mg, top, r, z);
  if (mg->null_space) project_mean(z, mg->L[top].lv.N);
  return rc;
}

static double dot(const double* a, const double* b, uint32_t n) {
  double s = 0.0;
  for (uint32_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* KSPSolve_CG (PETSc, left preconditioning), zero initial guess:
 *   r = b; z = B r; dp = ||z|| or ||r||; beta = z.r
 *   loop: p = z (+ beta/betaold p); w = A p; a = beta / p.w; x += a p; r -= a w;
 *         z = B r; dp = norm; test dp <= rtol dp0; beta = z.r */
int or_pcg(or_mg* mg, const double* b_in, double* x, double rtol, int max_it, int norm_type,
           double* history, int history_len, int* converged) {
  const int top = mg->nlevels - 1;
  const uint32_t N = mg->L[top].lv.N;
  double* b = (double*)malloc(sizeof(double) * N);
  double* r = (double*)malloc(sizeof(double) * N);
  double* z = (double*)malloc(sizeof(double) * N);
  double* p = (double*)malloc(sizeof(double) * N);
  double* w = (double*)malloc(sizeof(double) * N);
  memcpy(b, b_in, sizeof(double) * N);
  if (mg->null_space) project_mean(b, N);
  memset(x, 0, sizeof(double) * N);
  memcpy(r, b, sizeof(double) * N);
  or_mg_vcycle(mg, r, z);
  double dp = sqrt(norm_type ? dot(z, z, N) : dot(r, r, N));
  const double dp0 = dp;
  if (history && history_len > 0) history[0] = dp;
  double beta = dot(z, r, N), betaold = 0.0;
  int it = 0;
  *converged = (dp0 == 0.0);
  while (!*converged && it < max_it) {
    if (it == 0) {
      memcpy(p, z, sizeof(double) * N);
    } else {
      const double bb = beta / betaold;
      for (uint32_t i = 0; i < N; ++i) p[i] = z[i] + bb * p[i];
    }
    apply_A(mg, top, p, w);
    const double dpi = dot(p, w, N);
    betaold = beta;
    const double a = beta / dpi;
    for (uint32_t i = 0; i < N; ++i) x[i] += a * p[i];
    for (uint32_t i = 0; i < N; ++i) r[i] -= a * w[i];
    or_mg_vcycle(mg, r, z);
    dp = sqrt(norm_type ? dot(z, z, N) : dot(r, r, N));
    ++it;
    if (history && it < history_len) history[it] = dp;
    if (!(dp == dp)) break; /* NaN */
    if (dp <= rtol * dp0) *converged = 1;
    beta = dot(z, r, N);
  }
  free(b);
  free(r);
  free(z);
  free(p);
  free(w);
  return it;
}
