// Closed-form element stencils of the hanging-node trilinear element.
//
// An element of child number c with hanging mask H is the master cube [-1,1]^3 whose
// vertex k sits at its own corner when k is regular and at vertex k of the parent octant
// when k is hanging (the parent spans [-1,3] or [-3,1] along an axis, depending on the
// child bit on that axis).  Its 8 shape functions are the trilinear polynomials
// phi_i(x,y,z) = sum_m a_im B_m with B = {1, x, y, z, xy, yz, zx, xyz} and phi_i(X_k) = d_ik
// (Matlab/coarseShapeFunction3DCoeffs.m, Matlab/parentOfCoarse3d.m).
//   Laplacian  L[c][t][i][j] = int_{[-1,1]^3} grad phi_i . grad phi_j
//   mass       M[c][t][i][j] = int_{[-1,1]^3} phi_i phi_j
//   restriction Type2 R2[c][t][i][f]        = phi_i(vertex f of the element)
//   restriction Type1 R1[c][cf][t][i][f]    = phi_i(vertex f of child cf of the element)
// (Matlab/computeLapStencilsType2.m, computeMassStencilsType2.m, computeRPstencilsType{1,2}.m).
// The reference integrates with the 2-point Gauss rule; the integrands are products of two
// multilinear polynomials, so the rule is exact and this file uses the exact monomial
// moments int_{-1}^{1} t^0 = 2, t^1 = 0, t^2 = 2/3 instead.
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace dgh {

const uint8_t kEtypeMask[8][18] = {
    {0, 4, 2, 6, 16, 20, 18, 22, 14, 30, 84, 86, 94, 50, 54, 62, 118, 126},
    {0, 1, 8, 9, 32, 33, 40, 41, 13, 45, 49, 57, 61, 168, 169, 173, 185, 189},
    {0, 8, 1, 9, 64, 72, 65, 73, 11, 75, 200, 201, 203, 81, 89, 91, 217, 219},
    {0, 2, 4, 6, 128, 130, 132, 134, 7, 135, 162, 166, 167, 196, 198, 199, 230, 231},
    {0, 1, 32, 33, 64, 65, 96, 97, 35, 99, 69, 101, 103, 224, 225, 227, 229, 231},
    {0, 2, 128, 130, 16, 18, 144, 146, 138, 154, 19, 147, 155, 208, 210, 218, 211, 219},
    {0, 4, 16, 20, 128, 132, 144, 148, 21, 149, 140, 156, 157, 176, 180, 181, 188, 189},
    {0, 8, 64, 72, 32, 40, 96, 104, 76, 108, 42, 106, 110, 112, 120, 124, 122, 126},
};

int etype_of(int child, int mask) {
  if (child < 0 || child > 7) return -1;
  for (int t = 0; t < 18; ++t)
    if (kEtypeMask[child][t] == mask) return t;
  return -1;
}

namespace {

// exponents of the monomial basis B_m along x, y, z
constexpr int kExp[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                            {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};

double moment(int p) { return p == 0 ? 2.0 : (p == 1 ? 0.0 : 2.0 / 3.0); }

// Node position of vertex k of an element of child number c, hanging mask `mask`.
void node_position(int c, int mask, int k, double X[3]) {
  for (int d = 0; d < 3; ++d) {
    const int kb = (k >> d) & 1, cb = (c >> d) & 1;
    if (mask & (1 << k))
      X[d] = cb ? (kb ? 1.0 : -3.0) : (kb ? 3.0 : -1.0);   // parent's vertex k
    else
      X[d] = kb ? 1.0 : -1.0;
  }
}

// Shape-function coefficients a[i][m] (phi_i = sum_m a[i][m] B_m) by Gauss-Jordan on the
// 8x8 interpolation matrix V[k][m] = B_m(X_k): a = (V^{-1})^T.
void shape_coeffs(int c, int mask, double a[8][8]) {
  double V[8][16];
  for (int k = 0; k < 8; ++k) {
    double X[3];
    node_position(c, mask, k, X);
    for (int m = 0; m < 8; ++m) {
      double v = 1.0;
      for (int d = 0; d < 3; ++d)
        if (kExp[m][d]) v *= X[d];
      V[k][m] = v;
    }
    for (int m = 0; m < 8; ++m) V[k][8 + m] = (m == k) ? 1.0 : 0.0;
  }
  for (int col = 0; col < 8; ++col) {
    int piv = col;
    for (int r = col + 1; r < 8; ++r)
      if (std::fabs(V[r][col]) > std::fabs(V[piv][col])) piv = r;
    if (piv != col)
      for (int q = 0; q < 16; ++q) std::swap(V[col][q], V[piv][q]);
    const double inv = 1.0 / V[col][col];
    for (int q = 0; q < 16; ++q) V[col][q] *= inv;
    for (int r = 0; r < 8; ++r) {
      if (r == col) continue;
      const double f = V[r][col];
      if (f == 0.0) continue;
      for (int q = 0; q < 16; ++q) V[r][q] -= f * V[col][q];
    }
  }
  // V[:,8:] = Vinv with Vinv[m][k]: coefficient m of phi_k
  for (int i = 0; i < 8; ++i)
    for (int m = 0; m < 8; ++m) a[i][m] = V[m][8 + i];
}

// Shape-function values at points of the {-1,0,1} lattice are dyadic rationals; snap the
// rounding residue of the inversion so the restriction weights are exact.
double eval_phi(const double a[8], const double P[3]) {
  double v = 0.0;
  for (int m = 0; m < 8; ++m) {
    double b = a[m];
    for (int d = 0; d < 3; ++d)
      if (kExp[m][d]) b *= P[d];
    v += b;
  }
  const double r = std::nearbyint(v * 4096.0) / 4096.0;
  return (std::fabs(r - v) < 1e-12) ? r : v;
}

}  // namespace

void generate_stencils(double* lap2, double* mass2, double* rmat2, double* rmat1) {
  for (int c = 0; c < 8; ++c) {
    for (int t = 0; t < 18; ++t) {
      double a[8][8];
      shape_coeffs(c, kEtypeMask[c][t], a);
      const size_t base = (size_t(c) * 18 + t) * 64;
      for (int i = 0; i < 8; ++i) {
        for (int j = 0; j < 8; ++j) {
          double L = 0.0, M = 0.0;
          for (int m = 0; m < 8; ++m) {
            for (int n = 0; n < 8; ++n) {
              const double cc = a[i][m] * a[j][n];
              if (cc == 0.0) continue;
              const int ex = kExp[m][0] + kExp[n][0], ey = kExp[m][1] + kExp[n][1],
                        ez = kExp[m][2] + kExp[n][2];
              M += cc * moment(ex) * moment(ey) * moment(ez);
              // d/dx: both monomials must contain x; the derivative removes one power each
              if (kExp[m][0] && kExp[n][0]) L += cc * moment(ex - 2) * moment(ey) * moment(ez);
              if (kExp[m][1] && kExp[n][1]) L += cc * moment(ex) * moment(ey - 2) * moment(ez);
              if (kExp[m][2] && kExp[n][2]) L += cc * moment(ex) * moment(ey) * moment(ez - 2);
            }
          }
          if (lap2) lap2[base + i * 8 + j] = L;
          if (mass2) mass2[base + i * 8 + j] = M;
        }
      }
      if (rmat2) {
        for (int i = 0; i < 8; ++i)
          for (int f = 0; f < 8; ++f) {
            const double P[3] = {(f & 1) ? 1.0 : -1.0, (f & 2) ? 1.0 : -1.0, (f & 4) ? 1.0 : -1.0};
            rmat2[base + i * 8 + f] = eval_phi(a[i], P);
          }
      }
      if (rmat1) {
        for (int cf = 0; cf < 8; ++cf) {
          const size_t b1 = ((size_t(c) * 8 + cf) * 18 + t) * 64;
          for (int i = 0; i < 8; ++i)
            for (int f = 0; f < 8; ++f) {
              double P[3];
              for (int d = 0; d < 3; ++d) P[d] = -1.0 + ((cf >> d) & 1) + ((f >> d) & 1);
              rmat1[b1 + i * 8 + f] = eval_phi(a[i], P);
            }
        }
      }
    }
  }
}

// Grad-div stencils of the vector (3 dofs per node) element:
// GD[c][t][3k+a][3j+b] = int_{[-1,1]^3} d_a phi_k d_b phi_j (Matlab/computeGradDivType2.m; the
// reference reads them from Stencils/GDType2Stencils.inp in createGDmatType2,
// examples/src/backend/handleType2Stencils.C:303-330), by the exact monomial moments.
void generate_graddiv(double* GD /* [144][24][24] */) {
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double a[8][8];
      shape_coeffs(c, kEtypeMask[c][t], a);
      double* G = GD + (size_t(c) * 18 + t) * 576;
      for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 8; ++j)
          for (int da = 0; da < 3; ++da)
            for (int db = 0; db < 3; ++db) {
              double v = 0.0;
              for (int m = 0; m < 8; ++m) {
                if (!kExp[m][da]) continue;
                for (int n = 0; n < 8; ++n) {
                  if (!kExp[n][db]) continue;
                  const double cc = a[k][m] * a[j][n];
                  if (cc == 0.0) continue;
                  int e[3];
                  for (int d = 0; d < 3; ++d) e[d] = kExp[m][d] + kExp[n][d] - (d == da) - (d == db);
                  v += cc * moment(e[0]) * moment(e[1]) * moment(e[2]);
                }
              }
              G[(3 * k + da) * 24 + 3 * j + db] = v;
            }
    }
}

// Shape functions at the 3-point Gauss nodes of the element, in the order of
// examples/src/backend/omgRhs.C:1141-1180 (g = {0, +sqrt(3/5), -sqrt(3/5)}, m along x, n along
// y, p along z): S[s][j][9 m + 3 n + p] = phi_j(g_m, g_n, g_p), s = 18 c + t.  Plays the role of
// createShFnMat (examples/src/backend/handleType2Stencils.C:197-230), which reads
// Stencils/ShFnStencils.inp.
void generate_shfn(double* S /* [144][8][27] */) {
  const double g[3] = {0.0, std::sqrt(3.0 / 5.0), -std::sqrt(3.0 / 5.0)};
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double a[8][8];
      shape_coeffs(c, kEtypeMask[c][t], a);
      for (int j = 0; j < 8; ++j)
        for (int q = 0; q < 27; ++q) {
          const double P[3] = {g[q / 9], g[(q / 3) % 3], g[q % 3]};
          double v = 0.0;
          for (int m = 0; m < 8; ++m) {
            double b = a[j][m];
            for (int d = 0; d < 3; ++d)
              if (kExp[m][d]) b *= P[d];
            v += b;
          }
          S[((size_t(c) * 18 + t) * 8 + j) * 27 + q] = v;
        }
    }
}

// Weights of the 8 coarse shape functions of stencil (c,t) at the 27 positions
// {-1,0,1}^3 of the element, pos = tx + 3 ty + 9 tz: W[s][pos][k], s = 18 c + t.
void generate_transfer_weights(double* W /* [144][27][8] */) {
  for (int c = 0; c < 8; ++c)
    for (int t = 0; t < 18; ++t) {
      double a[8][8];
      shape_coeffs(c, kEtypeMask[c][t], a);
      for (int pos = 0; pos < 27; ++pos) {
        const double P[3] = {double(pos % 3) - 1.0, double((pos / 3) % 3) - 1.0,
                             double(pos / 9) - 1.0};
        for (int k = 0; k < 8; ++k) W[((size_t(c) * 18 + t) * 27 + pos) * 8 + k] = eval_phi(a[k], P);
      }
    }
}

}  // namespace dgh
