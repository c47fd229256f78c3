/*
 * learned.c -- CPU restatement of the learned Jacobi damping (per-element MLP -> nodal
 * damping weights) and of its training objective.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no learned smoother; its
 * smoother is damped Jacobi (bin/options:187-193), which is what every learned weight
 * falls back to.  Pinning: or_mlp_backward and or_learn_loss are checked against central
 * finite differences of or_mlp_forward / or_learn_loss (tests/test_learned_oracle.py);
 * the network with zero hidden weights reduces to constant weights, and the objective at
 * omega = 0 is exactly 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* round-to-nearest-even float -> bfloat16 -> float */
static float bf16r(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

/* parameter offsets: W1[NH][NF] b1[NH] W2[NH][NH] b2[NH] W3[NO][NH] b3[NO] */
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + OR_NH * OR_NF)
#define OFF_W2 (OFF_B1 + OR_NH)
#define OFF_B2 (OFF_W2 + OR_NH * OR_NH)
#define OFF_W3 (OFF_B2 + OR_NH)
#define OFF_B3 (OFF_W3 + OR_NO * OR_NH)

static void weights(const float* theta, int round, double* w) {
  for (int p = 0; p < OR_NPARAM; ++p) {
    const int is_bias = (p >= OFF_B1 && p < OFF_W2) || (p >= OFF_B2 && p < OFF_W3) || p >= OFF_B3;
    w[p] = (round && !is_bias) ? (double)bf16r(theta[p]) : (double)theta[p];
  }
}

void or_learn_node_logk(uint32_t E, uint32_t N, const uint32_t* nlist, const double* kappa, double lap_fac,
                        float* lkbar, float* lke) {
  double* s = calloc(N, sizeof(double));
  uint32_t* c = calloc(N, sizeof(uint32_t));
  for (uint32_t e = 0; e < E; ++e) {
    const double k = kappa ? kappa[e] : lap_fac;
    lke[e] = (float)log10(k);
    for (int v = 0; v < 8; ++v) {
      s[nlist[8 * (size_t)e + v]] += k;
      c[nlist[8 * (size_t)e + v]]++;
    }
  }
  for (uint32_t n = 0; n < N; ++n) lkbar[n] = c[n] ? (float)log10(s[n] / c[n]) : 0.0f;
  free(s);
  free(c);
}

void or_learn_features(uint32_t E, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                       const uint8_t* hnmask, const uint8_t* lev, const float* lkbar, const float* lke,
                       float* X) {
  for (uint32_t e = 0; e < E; ++e) {
    float* x = X + (size_t)e * OR_NF;
    memset(x, 0, sizeof(float) * OR_NF);
    x[0] = (float)((double)(lev[e] - 1) / (double)(max_depth - 1));
    x[1 + or_etype(child[e], hnmask[e])] = 1.0f;
    x[19 + child[e]] = 1.0f;
    for (int v = 0; v < 8; ++v) x[27 + v] = lkbar[nlist[8 * (size_t)e + v]] - lke[e];
  }
}

void or_mlp_forward(const float* theta, int round, uint32_t E, const float* X, double* O) {
  double* w = malloc(sizeof(double) * OR_NPARAM);
  weights(theta, round, w);
  for (uint32_t e = 0; e < E; ++e) {
    const float* x = X + (size_t)e * OR_NF;
    double h1[OR_NH], h2[OR_NH];
    for (int j = 0; j < OR_NH; ++j) {
      double a = w[OFF_B1 + j];
      for (int i = 0; i < OR_NF; ++i) a += w[OFF_W1 + j * OR_NF + i] * x[i];
      h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < OR_NH; ++j) {
      double a = w[OFF_B2 + j];
      for (int i = 0; i < OR_NH; ++i) a += w[OFF_W2 + j * OR_NH + i] * h1[i];
      h2[j] = a > 0 ? a : 0;
    }
    for (int m = 0; m < OR_NO; ++m) {
      double a = w[OFF_B3 + m];
      for (int j = 0; j < OR_NH; ++j) a += w[OFF_W3 + m * OR_NH + j] * h2[j];
      O[(size_t)e * OR_NO + m] = a;
    }
  }
  free(w);
}

/* grad[p] = sum_e sum_m dO[e][m] d o[e][m] / d theta_p (the bf16 rounding of the weights is
 * treated as the identity, as in straight-through training) */
void or_mlp_backward(const float* theta, int round, uint32_t E, const float* X, const double* dO, double* grad) {
  double* w = malloc(sizeof(double) * OR_NPARAM);
  weights(theta, round, w);
  memset(grad, 0, sizeof(double) * OR_NPARAM);
  for (uint32_t e = 0; e < E; ++e) {
    const float* x = X + (size_t)e * OR_NF;
    double h1[OR_NH], h2[OR_NH], d2[OR_NH], d1[OR_NH];
    for (int j = 0; j < OR_NH; ++j) {
      double a = w[OFF_B1 + j];
      for (int i = 0; i < OR_NF; ++i) a += w[OFF_W1 + j * OR_NF + i] * x[i];
      h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < OR_NH; ++j) {
      double a = w[OFF_B2 + j];
      for (int i = 0; i < OR_NH; ++i) a += w[OFF_W2 + j * OR_NH + i] * h1[i];
      h2[j] = a > 0 ? a : 0;
    }
    const double* go = dO + (size_t)e * OR_NO;
    for (int m = 0; m < OR_NO; ++m) {
      grad[OFF_B3 + m] += go[m];
      for (int j = 0; j < OR_NH; ++j) grad[OFF_W3 + m * OR_NH + j] += go[m] * h2[j];
    }
    for (int j = 0; j < OR_NH; ++j) {
      double a = 0;
      for (int m = 0; m < OR_NO; ++m) a += w[OFF_W3 + m * OR_NH + j] * go[m];
      d2[j] = h2[j] > 0 ? a : 0;
    }
    for (int j = 0; j < OR_NH; ++j) {
      grad[OFF_B2 + j] += d2[j];
      for (int i = 0; i < OR_NH; ++i) grad[OFF_W2 + j * OR_NH + i] += d2[j] * h1[i];
    }
    for (int i = 0; i < OR_NH; ++i) {
      double a = 0;
      for (int j = 0; j < OR_NH; ++j) a += w[OFF_W2 + j * OR_NH + i] * d2[j];
      d1[i] = h1[i] > 0 ? a : 0;
    }
    for (int j = 0; j < OR_NH; ++j) {
      grad[OFF_B1 + j] += d1[j];
      for (int i = 0; i < OR_NF; ++i) grad[OFF_W1 + j * OR_NF + i] += d1[j] * x[i];
    }
  }
  free(w);
}

/* cap_n = 2 A_nn / sum_{(e,k): nlist[e][k] = n} sum_j |K_e[k][j]| : any weights in (0, cap_n]
 * keep 2 D W^-1 - A diagonally dominant, so every weighted Jacobi sweep stays an A-norm
 * contraction and the symmetric V-cycle stays positive definite. */
void or_learn_cap(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                  const uint8_t* hnmask, const uint8_t* lev, const double* kappa, const double* mass,
                  double lap_fac, double mass_fac, const double* lap2, const double* mass2, double* cap) {
  double* d = calloc(N, sizeof(double));
  double* r = calloc(N, sizeof(double));
  const double hfac = 1.0 / (double)(1u << (max_depth - 1));
  for (uint32_t e = 0; e < E; ++e) {
    const double h = hfac * (double)(1u << (max_depth - lev[e]));
    const double k1 = kappa ? kappa[e] : lap_fac;
    const double k2 = kappa ? (mass ? mass[e] : 0.0) : mass_fac;
    const double f1 = k1 * h / 2.0, f2 = k2 * h * h * h / 8.0;
    const int s = 18 * child[e] + or_etype(child[e], hnmask[e]);
    const double* L = lap2 + (size_t)s * 64;
    const double* M = mass2 + (size_t)s * 64;
    for (int k = 0; k < 8; ++k) {
      const uint32_t n = nlist[8 * (size_t)e + k];
      for (int j = 0; j < 8; ++j) {
        const double a = f1 * L[k * 8 + j] + f2 * M[k * 8 + j];
        r[n] += fabs(a);
        if (nlist[8 * (size_t)e + j] == n) d[n] += a;
      }
    }
  }
  for (uint32_t n = 0; n < N; ++n) cap[n] = r[n] > 0 ? 2.0 * d[n] / r[n] : 1.0;
  free(d);
  free(r);
}

/* theta_n = mean of the element outputs at node n (over (e, k) with nlist[e][k] = n, in
 * (e, k) order); omega_n = cap_n * sigmoid(theta_n) */
void or_learn_nodal(uint32_t E, uint32_t N, const uint32_t* nlist, const double* O, const double* cap,
                    double* theta_n, double* omega) {
  uint32_t* c = calloc(N, sizeof(uint32_t));
  memset(theta_n, 0, sizeof(double) * N);
  for (uint32_t e = 0; e < E; ++e)
    for (int k = 0; k < 8; ++k) {
      theta_n[nlist[8 * (size_t)e + k]] += O[(size_t)e * 8 + k];
      c[nlist[8 * (size_t)e + k]]++;
    }
  for (uint32_t n = 0; n < N; ++n) {
    theta_n[n] = c[n] ? theta_n[n] / c[n] : 0.0;
    omega[n] = cap[n] / (1.0 + exp(-theta_n[n]));
  }
  free(c);
}

void or_learn_nodal_grad(uint32_t E, uint32_t N, const uint32_t* nlist, const double* theta_n, const double* cap,
                         const double* g_omega, double* dO) {
  uint32_t* c = calloc(N, sizeof(uint32_t));
  for (uint32_t e = 0; e < E; ++e)
    for (int k = 0; k < 8; ++k) c[nlist[8 * (size_t)e + k]]++;
  for (uint32_t e = 0; e < E; ++e)
    for (int k = 0; k < 8; ++k) {
      const uint32_t n = nlist[8 * (size_t)e + k];
      const double sg = 1.0 / (1.0 + exp(-theta_n[n]));
      dO[(size_t)e * 8 + k] = g_omega[n] * cap[n] * sg * (1.0 - sg) / c[n];
    }
  free(c);
}

/* Objective of one level: L = (1/S) sum_s |S_w^nu e_s|_A^2 / |e_s|_A^2 with
 * S_w = I - diag(omega) D^-1 A; g_omega (may be NULL) receives dL/d omega. */
double or_learn_loss(uint32_t E, uint32_t N, uint32_t max_depth, const uint32_t* nlist, const uint8_t* child,
                     const uint8_t* hnmask, const uint8_t* lev, const double* kappa, const double* mass,
                     double lap_fac, double mass_fac, const double* lap2, const double* mass2,
                     const double* omega, const double* invd, int nu, int S, const double* E0, double* g_omega) {
  double* es = malloc(sizeof(double) * N * (nu + 1));
  double* zs = malloc(sizeof(double) * N * (nu > 0 ? nu : 1));
  double* t = malloc(sizeof(double) * N);
  double* g = malloc(sizeof(double) * N);
  double* u = malloc(sizeof(double) * N);
  double L = 0.0;
  if (g_omega) memset(g_omega, 0, sizeof(double) * N);
#define MV(x, y) or_matvec(E, N, max_depth, nlist, child, hnmask, lev, kappa, mass, lap_fac, mass_fac, lap2, mass2, x, y)
  for (int s = 0; s < S; ++s) {
    for (uint32_t n = 0; n < N; ++n) es[n] = E0[(size_t)n * S + s];
    MV(es, t);
    double ea = 0.0;
    for (uint32_t n = 0; n < N; ++n) ea += es[n] * t[n];
    for (int k = 0; k < nu; ++k) {
      double* ek = es + (size_t)k * N;
      double* en = es + (size_t)(k + 1) * N;
      double* z = zs + (size_t)k * N;
      if (k > 0) MV(ek, t);
      for (uint32_t n = 0; n < N; ++n) {
        z[n] = invd[n] * t[n];
        en[n] = ek[n] - omega[n] * z[n];
      }
    }
    double* ef = es + (size_t)nu * N;
    MV(ef, t);
    double fa = 0.0;
    for (uint32_t n = 0; n < N; ++n) fa += ef[n] * t[n];
    L += fa / ea;
    if (g_omega) {
      for (uint32_t n = 0; n < N; ++n) g[n] = 2.0 * t[n] / (ea * S);
      for (int k = nu - 1; k >= 0; --k) {
        const double* z = zs + (size_t)k * N;
        for (uint32_t n = 0; n < N; ++n) {
          g_omega[n] += -g[n] * z[n];
          u[n] = invd[n] * (omega[n] * g[n]);
        }
        if (k > 0) {
          MV(u, t);
          for (uint32_t n = 0; n < N; ++n) g[n] -= t[n];
        }
      }
    }
  }
#undef MV
  free(es);
  free(zs);
  free(t);
  free(g);
  free(u);
  return L / S;
}
