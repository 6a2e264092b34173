/*
 * adrt_oracle.c -- FP64 CPU restatement of the AddingDoublingRT column solver.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and reported CPU baseline).  The
 * product library never links, calls or falls back to this file.
 *
 * Restated from the reference C++ solver (paths relative to the reference
 * repository):
 *   src/quadrature.cpp:11-76      Gauss-Legendre nodes, Legendre recurrence
 *   src/phase_matrix.hpp:61-181   P++/P+- and solar phase vectors, Hansen scaling
 *   src/planck.cpp:35-177         band Planck function and dB/dT
 *   src/doubling.hpp:20-230       thin-layer init + doubling recursion
 *   src/adding.hpp:31-258         adding of two layers (3 variants + dispatch)
 *   src/solver.cpp:24-73          interface intensities
 *   src/solver.cpp:117-341        analytic temperature Jacobian
 *   src/solver.cpp:348-722        column solve (delta-M, surface, sweeps, outputs)
 *   src/matrix.hpp:120-124        right-solve X = B A^-1 via LU(A^T)
 *   src/adding_doubling.hpp:123-174  validation
 * Dense algebra (the reference uses Eigen 3.4) is restated with an unblocked
 * partial-pivot LU that picks pivots the way Eigen's PartialPivLU does.
 */
#include "adrt_oracle.h"

/* Internal arithmetic type: double for the oracle proper; long double (x87,
 * 64-bit mantissa) for the extended-precision build that measures how far an
 * FP64 solve sits from the exact solution of the same discrete problem. */
#ifdef ORC_LONG_DOUBLE
typedef long double real_t;
#define ORC_MAXR LDBL_MAX
#define ORC_EPSR DBL_EPSILON
#else
typedef double real_t;
#define ORC_MAXR DBL_MAX
#define ORC_EPSR DBL_EPSILON
#endif

#include <float.h>
#include <tgmath.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const real_t PI = 3.14159265358979323846264338327950288L; /* src/constants.hpp:8 */

/* ------------------------------------------------------------------------ */
/* error handling                                                            */
/* ------------------------------------------------------------------------ */

static int set_err(char* err, int errlen, int code, const char* msg) {
  if (err && errlen > 0) {
    strncpy(err, msg, (size_t)errlen - 1);
    err[errlen - 1] = '\0';
  }
  return code;
}

/* ------------------------------------------------------------------------ */
/* quadrature (src/quadrature.cpp:11-76)                                     */
/* ------------------------------------------------------------------------ */

static void gl_r(int n, real_t* nodes, real_t* weights) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    real_t z = cos(PI * (i + 0.75) / (n + 0.5));
    real_t pp = 0.0;
    for (int iter = 0; iter < 100; ++iter) {
      real_t p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        real_t pk = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      pp = n * (z * p1 - p0) / (z * z - 1.0);
      real_t dz = -p1 / pp;
      z += dz;
      if (fabs(dz) < 1e-15) break;
    }
    real_t w = 2.0 / ((1.0 - z * z) * pp * pp);
    nodes[i] = (1.0 - z) / 2.0;
    nodes[n - 1 - i] = (1.0 + z) / 2.0;
    weights[i] = w / 2.0;
    weights[n - 1 - i] = w / 2.0;
  }
}

static void legtab_r(int L, const real_t* x, int nx, real_t* Pl) {
  for (int i = 0; i < nx; ++i) Pl[i] = 1.0;
  if (L > 1)
    for (int i = 0; i < nx; ++i) Pl[nx + i] = x[i];
  for (int l = 2; l < L; ++l)
    for (int i = 0; i < nx; ++i)
      Pl[l * nx + i] =
          ((2 * l - 1) * x[i] * Pl[(l - 1) * nx + i] - (l - 1) * Pl[(l - 2) * nx + i]) / l;
}

/* ------------------------------------------------------------------------ */
/* phase matrices (src/phase_matrix.hpp:61-111, 130-181)                     */
/* ------------------------------------------------------------------------ */

static void phase_from_table(const real_t* chi, int L, const real_t* Pl, int n,
                             const real_t* wt, real_t* Ppp, real_t* Ppm) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      real_t sum_pp = 0.0, sum_pm = 0.0, sign = 1.0;
      for (int l = 0; l < L; ++l) {
        real_t term = (2 * l + 1) * chi[l] * Pl[l * n + i] * Pl[l * n + j];
        sum_pp += term;
        sum_pm += sign * term;
        sign = -sign;
      }
      Ppp[i * n + j] = sum_pp / (2.0 * PI);
      Ppm[i * n + j] = sum_pm / (2.0 * PI);
    }
  for (int j = 0; j < n; ++j) {
    real_t sum = 0.0;
    for (int i = 0; i < n; ++i) sum += (Ppp[i * n + j] + Ppm[i * n + j]) * wt[i];
    if (sum > 0.0) {
      real_t correction = 1.0 / (2.0 * PI * sum);
      for (int i = 0; i < n; ++i) {
        Ppp[i * n + j] *= correction;
        Ppm[i * n + j] *= correction;
      }
    }
  }
}

static void solar_from_table(const real_t* chi, int L, const real_t* Pl, int n,
                             const real_t* wt, real_t mu0, real_t* pp, real_t* pm) {
  real_t* Plmu0 = (real_t*)malloc(sizeof(real_t) * (size_t)(L > 1 ? L : 2));
  Plmu0[0] = 1.0;
  if (L > 1) Plmu0[1] = mu0;
  for (int l = 2; l < L; ++l)
    Plmu0[l] = ((2 * l - 1) * mu0 * Plmu0[l - 1] - (l - 1) * Plmu0[l - 2]) / l;
  for (int i = 0; i < n; ++i) {
    real_t sum_p = 0.0, sum_m = 0.0, sign = 1.0;
    for (int l = 0; l < L; ++l) {
      real_t term = (2 * l + 1) * chi[l] * Pl[l * n + i] * Plmu0[l];
      sum_p += term;
      sum_m += sign * term;
      sign = -sign;
    }
    pp[i] = sum_p / (2.0 * PI);
    pm[i] = sum_m / (2.0 * PI);
  }
  real_t sum = 0.0;
  for (int i = 0; i < n; ++i) sum += (pp[i] + pm[i]) * wt[i];
  if (sum > 0.0) {
    real_t correction = 1.0 / (2.0 * PI * sum);
    for (int i = 0; i < n; ++i) {
      pp[i] *= correction;
      pm[i] *= correction;
    }
  }
  free(Plmu0);
}

static void phase_r(const real_t* chi, int L, const real_t* mu, const real_t* wt,
                        int n, real_t* Ppp, real_t* Ppm) {
  real_t* Pl = (real_t*)malloc(sizeof(real_t) * (size_t)(L * n + 1));
  legtab_r(L, mu, n, Pl);
  phase_from_table(chi, L, Pl, n, wt, Ppp, Ppm);
  free(Pl);
}

static void solar_r(const real_t* chi, int L, const real_t* mu,
                             const real_t* wt, int n, real_t mu0, real_t* p_plus,
                             real_t* p_minus) {
  real_t* Pl = (real_t*)malloc(sizeof(real_t) * (size_t)(L * n + 1));
  legtab_r(L, mu, n, Pl);
  solar_from_table(chi, L, Pl, n, wt, mu0, p_plus, p_minus);
  free(Pl);
}

/* ------------------------------------------------------------------------ */
/* Planck (src/planck.cpp:13-177)                                            */
/* ------------------------------------------------------------------------ */

static const real_t P_C2 = 1.438786, P_SIGMA = 5.67032E-8, P_VCUT = 1.5;
static const real_t P_A1 = 1.0 / 3.0, P_A2 = -1.0 / 8.0, P_A3 = 1.0 / 60.0,
                    P_A4 = -1.0 / 5040.0, P_A5 = 1.0 / 272160.0,
                    P_A6 = -1.0 / 13305600.0;

static real_t planck_kernel(real_t v) { return v * v * v / (exp(v) - 1.0); }

static int planck_r(real_t wnumlo, real_t wnumhi, real_t temp, real_t* out) {
  if (temp < 0.0 || wnumhi < wnumlo || wnumlo < 0.0) return 16;
  if (temp < 1.e-4) { *out = 0.0; return 0; }
  if (wnumhi == wnumlo) {
    real_t arg = exp(-P_C2 * wnumhi / temp);
    *out = 1.1911e-8 * wnumhi * wnumhi * wnumhi * arg / (1.0 - arg);
    return 0;
  }
  real_t vmax = log((real_t)DBL_MAX);
  real_t sigdpi = P_SIGMA / PI;
  real_t conc = 15.0 / pow(PI, 4.0);
  real_t v[2] = {P_C2 * wnumlo / temp, P_C2 * wnumhi / temp};
  if (v[0] > ORC_EPSR && v[1] < vmax && (wnumhi - wnumlo) / wnumhi < 1.e-2) {
    real_t hh = v[1] - v[0];
    real_t oldval = 0.0;
    real_t val0 = planck_kernel(v[0]) + planck_kernel(v[1]);
    for (int n = 1; n <= 10; ++n) {
      real_t del = hh / (2.0 * n);
      real_t val = val0;
      for (int k = 1; k <= 2 * n - 1; ++k)
        val += 2.0 * (1 + k % 2) * planck_kernel(v[0] + k * del);
      val *= del * P_A1;
      if (fabs((val - oldval) / val) <= 1.e-6) {
        *out = sigdpi * pow(temp, 4.0) * conc * val;
        return 0;
      }
      oldval = val;
    }
    *out = sigdpi * pow(temp, 4.0) * conc * oldval;
    return 0;
  }
  real_t p[2] = {0.0, 0.0}, d[2] = {0.0, 0.0};
  int smallv = 0;
  static const real_t vcp[7] = {10.25, 5.7, 3.9, 2.9, 2.3, 1.9, 0.0};
  for (int i = 0; i < 2; ++i) {
    if (v[i] < P_VCUT) {
      smallv++;
      real_t vsq = v[i] * v[i];
      p[i] = conc * vsq * v[i] *
             (P_A1 + v[i] * (P_A2 + v[i] * (P_A3 + vsq * (P_A4 + vsq * (P_A5 + vsq * P_A6)))));
    } else {
      int mmax = 1;
      while (v[i] < vcp[mmax - 1]) mmax++;
      real_t ex = exp(-v[i]);
      real_t exm = 1.0;
      d[i] = 0.0;
      for (int m = 1; m <= mmax; ++m) {
        real_t mv = m * v[i];
        exm *= ex;
        d[i] += exm * (6.0 + mv * (6.0 + mv * (3.0 + mv))) / (m * m * m * m);
      }
      d[i] *= conc;
    }
  }
  real_t ans;
  if (smallv == 2) ans = p[1] - p[0];
  else if (smallv == 1) ans = 1.0 - p[0] - d[1];
  else ans = d[0] - d[1];
  ans *= sigdpi * pow(temp, 4.0);
  *out = (ans == 0.0) ? 0.0 : ans;
  return 0;
}

static real_t planck_fkernel(real_t v) {
  if (v < 1.e-6) return v * v;
  real_t em = exp(-v);
  return v * v * v * em / (1.0 - em);
}

static int planckd_r(real_t wnumlo, real_t wnumhi, real_t temp, real_t* out) {
  if (temp < 0.0 || wnumhi < wnumlo || wnumlo < 0.0) return 16;
  if (temp < 1.e-4) { *out = 0.0; return 0; }
  if (wnumhi == wnumlo) {
    real_t x = P_C2 * wnumhi / temp;
    real_t em = exp(-x);
    real_t b = 1.1911e-8 * wnumhi * wnumhi * wnumhi * em / (1.0 - em);
    *out = b * (x / temp) / (1.0 - em);
    return 0;
  }
  real_t sigdpi = P_SIGMA / PI;
  real_t conc = 15.0 / pow(PI, 4.0);
  real_t v0 = P_C2 * wnumlo / temp, v1 = P_C2 * wnumhi / temp;
  real_t B;
  planck_r(wnumlo, wnumhi, temp, &B);
  real_t boundary = sigdpi * conc * temp * temp * temp *
                    (v1 * planck_fkernel(v1) - v0 * planck_fkernel(v0));
  *out = 4.0 * B / temp - boundary;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* dense algebra (restates the Eigen calls in src/matrix.hpp:65-124)         */
/* ------------------------------------------------------------------------ */

/* C = A*B (n x n, row-major); C must not alias A or B. */
static void mm(int n, const real_t* A, const real_t* B, real_t* C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      real_t s = 0.0;
      for (int k = 0; k < n; ++k) s += A[i * n + k] * B[k * n + j];
      C[i * n + j] = s;
    }
}

static void mv(int n, const real_t* A, const real_t* x, real_t* y) {
  for (int i = 0; i < n; ++i) {
    real_t s = 0.0;
    for (int k = 0; k < n; ++k) s += A[i * n + k] * x[k];
    y[i] = s;
  }
}

/* In-place LU with partial (row) pivoting: pivot = first largest |a_ik|,
 * multipliers formed by division, as Eigen's unblocked PartialPivLU. */
static void lu_factor(int n, real_t* A, int* perm) {
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    int p = k;
    real_t big = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      real_t v = fabs(A[i * n + k]);
      if (v > big) { big = v; p = i; }
    }
    if (big != 0.0) {
      if (p != k) {
        for (int j = 0; j < n; ++j) {
          real_t t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t;
        }
        int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
      }
      real_t piv = A[k * n + k];
      for (int i = k + 1; i < n; ++i) A[i * n + k] /= piv;
    }
    for (int i = k + 1; i < n; ++i) {
      real_t lik = A[i * n + k];
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= lik * A[k * n + j];
    }
  }
}

/* Solve (P^T L U) x = b given lu_factor output; x may alias nothing. */
static void lu_solve(int n, const real_t* LU, const int* perm, const real_t* b, real_t* x) {
  for (int i = 0; i < n; ++i) x[i] = b[perm[i]];
  for (int i = 0; i < n; ++i) {
    real_t s = x[i];
    for (int k = 0; k < i; ++k) s -= LU[i * n + k] * x[k];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    real_t s = x[i];
    for (int k = i + 1; k < n; ++k) s -= LU[i * n + k] * x[k];
    x[i] = s / LU[i * n + i];
  }
}

/* Solve A x = b (src/matrix.hpp:82-84). */
static void solve_vec(int n, const real_t* A, const real_t* b, real_t* x, real_t* work,
                      int* perm) {
  memcpy(work, A, sizeof(real_t) * (size_t)(n * n));
  lu_factor(n, work, perm);
  lu_solve(n, work, perm, b, x);
}

/* X = B A^-1 via LU(A^T), solving A^T x_r = b_r for each row r of B
 * (src/matrix.hpp:120-124).  X must not alias A or B. */
static void right_solve(int n, const real_t* A, const real_t* B, real_t* X, real_t* work,
                        int* perm, real_t* tmp) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) work[i * n + j] = A[j * n + i];
  lu_factor(n, work, perm);
  for (int r = 0; r < n; ++r) lu_solve(n, work, perm, B + r * n, X + r * n);
  (void)tmp;
}

/* ------------------------------------------------------------------------ */
/* layer store (src/layer.hpp:19-49)                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
  real_t *Rab, *Rba, *Tab, *Tba;
  real_t *su, *sd, *sus, *sds, *jp, *jq;
  int scat;
} layer_t;

static size_t layer_doubles(int n) { return (size_t)(4 * n * n + 6 * n); }

static void layer_bind(layer_t* L, real_t* mem, int n) {
  L->Rab = mem;
  L->Rba = mem + n * n;
  L->Tab = mem + 2 * n * n;
  L->Tba = mem + 3 * n * n;
  L->su = mem + 4 * n * n;
  L->sd = L->su + n;
  L->sus = L->sd + n;
  L->sds = L->sus + n;
  L->jp = L->sds + n;
  L->jq = L->jp + n;
  L->scat = 0;
}

/* default LayerMatrices: R = 0, T = I, sources 0, not scattering */
static void layer_default(layer_t* L, int n) {
  memset(L->Rab, 0, sizeof(real_t) * layer_doubles(n));
  for (int i = 0; i < n; ++i) {
    L->Tab[i * n + i] = 1.0;
    L->Tba[i * n + i] = 1.0;
  }
  L->scat = 0;
}

/* ------------------------------------------------------------------------ */
/* doubling (src/doubling.hpp:20-230)                                        */
/* ------------------------------------------------------------------------ */

static int ipow0(real_t omega) {           /* src/doubling.hpp:20-26 */
  if (omega < 0.01) return 4;
  if (omega < 0.1) return 10;
  return 16;
}

static int nn_r(real_t tau, real_t omega) { /* src/doubling.hpp:81,107-109 */
  if (omega < 0.0) omega = 0.0;
  if (omega > 1.0) omega = 1.0;
  int nn = (int)(log(tau) / log(2.0)) + ipow0(omega);
  if (nn < 1) nn = 1;
  return nn;
}

static void doubling_layer(int n, real_t tau, real_t omega, real_t B_top, real_t B_bottom,
                           const real_t* Ppp, const real_t* Ppm, const real_t* mu,
                           const real_t* wt, real_t solar_flux, real_t solar_mu,
                           real_t tau_cumulative, const real_t* pps, const real_t* pms,
                           layer_t* out) {
  layer_default(out, n);
  real_t B_bar = (B_bottom + B_top) / 2.0;
  real_t B_d = (tau > 0.0) ? (B_bottom - B_top) / tau : 0.0;
  if (tau <= 0.0) return;

  if (omega <= 0.0) { /* pure absorption, src/doubling.hpp:56-77 */
    memset(out->Tab, 0, sizeof(real_t) * (size_t)(n * n));
    memset(out->Tba, 0, sizeof(real_t) * (size_t)(n * n));
    for (int i = 0; i < n; ++i) {
      real_t tex = -tau / mu[i];
      real_t trans = (tex > -200.0) ? exp(tex) : 0.0;
      out->Tab[i * n + i] = trans;
      out->Tba[i * n + i] = trans;
      real_t one_minus_t = 1.0 - trans;
      real_t slope_term = mu[i] * one_minus_t - 0.5 * tau * (1.0 + trans);
      out->su[i] = B_bar * one_minus_t + B_d * slope_term;
      out->sd[i] = B_bar * one_minus_t - B_d * slope_term;
      out->jp[i] = 0.5 * one_minus_t - slope_term / tau;
      out->jq[i] = 0.5 * one_minus_t + slope_term / tau;
    }
    return;
  }

  out->scat = 1;
  if (omega < 0.0) omega = 0.0;
  if (omega > 1.0) omega = 1.0;
  real_t con = 2.0 * omega * PI;

  size_t nn2 = (size_t)(n * n);
  real_t* buf = (real_t*)malloc(sizeof(real_t) * (nn2 * 10 + (size_t)n * 16));
  real_t *Gpp = buf, *Gpm = buf + nn2, *R = buf + 2 * nn2, *T = buf + 3 * nn2;
  real_t *Rsq = buf + 4 * nn2, *TG = buf + 5 * nn2, *TGR = buf + 6 * nn2;
  real_t *Rn = buf + 7 * nn2, *Tn = buf + 8 * nn2, *work = buf + 9 * nn2;
  real_t* v = buf + 10 * nn2;
  real_t *y = v, *z = v + n, *sus = v + 2 * n, *sds = v + 3 * n, *zpgy = v + 4 * n;
  real_t *a1 = v + 5 * n, *a2 = v + 6 * n, *a3 = v + 7 * n, *a4 = v + 8 * n;
  real_t *rup = v + 9 * n, *rdn = v + 10 * n, *ynew = v + 11 * n, *znew = v + 12 * n;
  real_t *susn = v + 13 * n, *sdsn = v + 14 * n;
  int perm[ORC_MAX_N];

  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      real_t pppc = Ppp[i * n + j] * wt[j];
      real_t temp = ((i == j) ? 1.0 : 0.0) + (-con) * pppc;
      Gpp[i * n + j] = temp / mu[i];
      real_t ppmc = Ppm[i * n + j] * wt[j];
      Gpm[i * n + j] = con * ppmc / mu[i];
    }

  int nn = (int)(log(tau) / log(2.0)) + ipow0(omega);
  if (nn < 1) nn = 1;
  real_t xfac = 1.0 / pow(2.0, nn);
  real_t tau0 = tau * xfac;

  int has_solar = (solar_flux > 0.0 && solar_mu > 0.0 && pps != NULL && pms != NULL);
  real_t F_top = has_solar ? solar_flux * exp(-tau_cumulative / solar_mu) : 0.0;

  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      T[i * n + j] = ((i == j) ? 1.0 : 0.0) - tau0 * Gpp[i * n + j];
      R[i * n + j] = tau0 * Gpm[i * n + j];
    }
  for (int i = 0; i < n; ++i) {
    y[i] = (1.0 - omega) * tau0 / mu[i];
    z[i] = 0.0;
    sus[i] = 0.0;
    sds[i] = 0.0;
  }
  if (has_solar)
    for (int i = 0; i < n; ++i) {
      real_t base = omega * tau0 / mu[i] * F_top;
      sus[i] = base * pms[i];
      sds[i] = base * pps[i];
    }
  real_t g = 0.5 * tau0;
  real_t gamma = has_solar ? exp(-tau0 / solar_mu) : 0.0;

  for (int k = 0; k < nn; ++k) {
    mm(n, R, R, Rsq);
    for (size_t e = 0; e < nn2; ++e) Rsq[e] = -Rsq[e];
    for (int i = 0; i < n; ++i) Rsq[i * n + i] += 1.0; /* I - R^2 */
    right_solve(n, Rsq, T, TG, work, perm, NULL);
    mm(n, TG, R, TGR);
    mm(n, TGR, T, Rn);
    for (size_t e = 0; e < nn2; ++e) Rn[e] = R[e] + Rn[e];
    mm(n, TG, T, Tn);

    for (int i = 0; i < n; ++i) zpgy[i] = z[i] + g * y[i];
    mv(n, TG, zpgy, a1);
    mv(n, TGR, zpgy, a2);
    mv(n, TG, y, a3);
    mv(n, TGR, y, a4);
    for (int i = 0; i < n; ++i) {
      znew[i] = (a1[i] - a2[i]) + z[i] - g * y[i];
      ynew[i] = a3[i] + a4[i] + y[i];
    }
    if (has_solar) {
      mv(n, R, sds, a1);
      mv(n, R, sus, a2);
      for (int i = 0; i < n; ++i) {
        rup[i] = a1[i] + gamma * sus[i];
        rdn[i] = gamma * a2[i] + sds[i];
      }
      mv(n, TG, rup, a3);
      mv(n, TG, rdn, a4);
      for (int i = 0; i < n; ++i) {
        susn[i] = a3[i] + sus[i];
        sdsn[i] = a4[i] + gamma * sds[i];
      }
      gamma = gamma * gamma;
      memcpy(sus, susn, sizeof(real_t) * (size_t)n);
      memcpy(sds, sdsn, sizeof(real_t) * (size_t)n);
    }
    memcpy(R, Rn, sizeof(real_t) * nn2);
    memcpy(T, Tn, sizeof(real_t) * nn2);
    memcpy(y, ynew, sizeof(real_t) * (size_t)n);
    memcpy(z, znew, sizeof(real_t) * (size_t)n);
    g = 2.0 * g;
  }

  memcpy(out->Rab, R, sizeof(real_t) * nn2);
  memcpy(out->Rba, R, sizeof(real_t) * nn2);
  memcpy(out->Tab, T, sizeof(real_t) * nn2);
  memcpy(out->Tba, T, sizeof(real_t) * nn2);
  for (int i = 0; i < n; ++i) {
    out->su[i] = y[i] * B_bar + z[i] * B_d;
    out->sd[i] = y[i] * B_bar - z[i] * B_d;
    out->sus[i] = sus[i];
    out->sds[i] = sds[i];
    out->jp[i] = 0.5 * y[i] - z[i] / tau;
    out->jq[i] = 0.5 * y[i] + z[i] / tau;
  }
  free(buf);
}

static int doubling_r(int n, real_t tau, real_t omega, real_t B_top, real_t B_bottom,
                 const real_t* Ppp, const real_t* Ppm, const real_t* mu, const real_t* wt,
                 real_t solar_flux, real_t solar_mu, real_t tau_cumulative,
                 const real_t* p_plus_solar, const real_t* p_minus_solar, real_t* R,
                 real_t* T, real_t* s_up, real_t* s_down, real_t* s_up_solar,
                 real_t* s_down_solar, real_t* j_p, real_t* j_q) {
  real_t* mem = (real_t*)malloc(sizeof(real_t) * layer_doubles(n));
  layer_t L;
  layer_bind(&L, mem, n);
  doubling_layer(n, tau, omega, B_top, B_bottom, Ppp, Ppm, mu, wt, solar_flux, solar_mu,
                 tau_cumulative, p_plus_solar, p_minus_solar, &L);
  size_t nn2 = (size_t)(n * n);
  memcpy(R, L.Rab, sizeof(real_t) * nn2);
  memcpy(T, L.Tab, sizeof(real_t) * nn2);
  memcpy(s_up, L.su, sizeof(real_t) * (size_t)n);
  memcpy(s_down, L.sd, sizeof(real_t) * (size_t)n);
  memcpy(s_up_solar, L.sus, sizeof(real_t) * (size_t)n);
  memcpy(s_down_solar, L.sds, sizeof(real_t) * (size_t)n);
  memcpy(j_p, L.jp, sizeof(real_t) * (size_t)n);
  memcpy(j_q, L.jq, sizeof(real_t) * (size_t)n);
  int s = L.scat;
  free(mem);
  return s;
}

/* ------------------------------------------------------------------------ */
/* adding (src/adding.hpp:19-258)                                            */
/* ------------------------------------------------------------------------ */

/* LU factors captured for the temperature Jacobian (src/adding.hpp:19-25). */
typedef struct {
  int general;
  real_t* lu1; /* of A1 = I - bot.R_ab top.R_ba */
  real_t* lu2; /* of A2 = I - top.R_ba bot.R_ab */
  int p1[ORC_MAX_N], p2[ORC_MAX_N];
} addops_t;

typedef struct {
  real_t *t1, *t2, *t3, *t4, *t5, *t6, *work, *v1, *v2, *v3;
} scratch_t;

static void scratch_alloc(scratch_t* s, int n, real_t** block) {
  size_t nn2 = (size_t)(n * n);
  *block = (real_t*)malloc(sizeof(real_t) * (nn2 * 7 + (size_t)n * 3));
  s->t1 = *block;
  s->t2 = s->t1 + nn2;
  s->t3 = s->t2 + nn2;
  s->t4 = s->t3 + nn2;
  s->t5 = s->t4 + nn2;
  s->t6 = s->t5 + nn2;
  s->work = s->t6 + nn2;
  s->v1 = s->work + nn2;
  s->v2 = s->v1 + n;
  s->v3 = s->v2 + n;
}

/* src/adding.hpp:31-62 */
static void add_sources(int n, layer_t* ans, const layer_t* top, const layer_t* bot,
                        const real_t* TbaD1, const real_t* TbcD2, scratch_t* s) {
  for (int pass = 0; pass < 2; ++pass) {
    const real_t* s1u = pass ? top->sus : top->su;
    const real_t* s1d = pass ? top->sds : top->sd;
    const real_t* s2u = pass ? bot->sus : bot->su;
    const real_t* s2d = pass ? bot->sds : bot->sd;
    real_t* au = pass ? ans->sus : ans->su;
    real_t* ad = pass ? ans->sds : ans->sd;
    mv(n, bot->Rab, s1d, s->v1);
    for (int i = 0; i < n; ++i) s->v2[i] = s2u[i] + s->v1[i];
    mv(n, TbaD1, s->v2, s->v3);
    for (int i = 0; i < n; ++i) au[i] = s1u[i] + s->v3[i];
    mv(n, top->Rba, s2u, s->v1);
    for (int i = 0; i < n; ++i) s->v2[i] = s1d[i] + s->v1[i];
    mv(n, TbcD2, s->v2, s->v3);
    for (int i = 0; i < n; ++i) ad[i] = s2d[i] + s->v3[i];
  }
}

/* src/adding.hpp:69-110 */
static void add_general(int n, const layer_t* top, const layer_t* bot, layer_t* ans,
                        addops_t* ops, scratch_t* s) {
  size_t nn2 = (size_t)(n * n);
  int perm[ORC_MAX_N];
  real_t *A1 = s->t1, *A2 = s->t2, *TbaD1 = s->t3, *TbcD2 = s->t4;
  mm(n, bot->Rab, top->Rba, A1);
  mm(n, top->Rba, bot->Rab, A2);
  for (size_t e = 0; e < nn2; ++e) { A1[e] = -A1[e]; A2[e] = -A2[e]; }
  for (int i = 0; i < n; ++i) { A1[i * n + i] += 1.0; A2[i * n + i] += 1.0; }
  if (ops) {
    /* T_ba_D1 = top.T_ba * inv(A1) with the LU of A1 kept for the Jacobian
     * (src/adding.hpp:318-325). */
    ops->general = 1;
    memcpy(ops->lu1, A1, sizeof(real_t) * nn2);
    memcpy(ops->lu2, A2, sizeof(real_t) * nn2);
    lu_factor(n, ops->lu1, ops->p1);
    lu_factor(n, ops->lu2, ops->p2);
    real_t* inv = s->t5;
    real_t* e = s->v1;
    for (int pass = 0; pass < 2; ++pass) {
      const real_t* LU = pass ? ops->lu2 : ops->lu1;
      const int* P = pass ? ops->p2 : ops->p1;
      for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) e[i] = (i == c) ? 1.0 : 0.0;
        lu_solve(n, LU, P, e, s->v2);
        for (int i = 0; i < n; ++i) inv[i * n + c] = s->v2[i];
      }
      mm(n, pass ? bot->Tab : top->Tba, inv, pass ? TbcD2 : TbaD1);
    }
  } else {
    right_solve(n, A1, top->Tba, TbaD1, s->work, perm, NULL);
    right_solve(n, A2, bot->Tab, TbcD2, s->work, perm, NULL);
  }
  mm(n, TbaD1, bot->Rab, s->t5);
  mm(n, s->t5, top->Tab, s->t6);
  for (size_t e = 0; e < nn2; ++e) ans->Rab[e] = top->Rab[e] + s->t6[e];
  mm(n, TbcD2, top->Rba, s->t5);
  mm(n, s->t5, bot->Tba, s->t6);
  for (size_t e = 0; e < nn2; ++e) ans->Rba[e] = bot->Rba[e] + s->t6[e];
  mm(n, TbcD2, top->Tab, ans->Tab);
  mm(n, TbaD1, bot->Tba, ans->Tba);
  add_sources(n, ans, top, bot, TbaD1, TbcD2, s);
  ans->scat = 1;
}

/* src/adding.hpp:117-174 */
static void add_nonscat_top(int n, const layer_t* top, const layer_t* bot, layer_t* ans,
                            addops_t* ops, scratch_t* s) {
  if (ops) ops->general = 0;
  real_t* t = s->v3;
  for (int i = 0; i < n; ++i) t[i] = top->Tab[i * n + i];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      ans->Rab[i * n + j] = t[i] * bot->Rab[i * n + j] * t[j];
      ans->Rba[i * n + j] = bot->Rba[i * n + j];
      ans->Tab[i * n + j] = bot->Tab[i * n + j] * t[j];
      ans->Tba[i * n + j] = t[i] * bot->Tba[i * n + j];
    }
  real_t* tmp = s->t6; /* vector scratch */
  for (int pass = 0; pass < 2; ++pass) {
    const real_t* s1u = pass ? top->sus : top->su;
    const real_t* s1d = pass ? top->sds : top->sd;
    const real_t* s2u = pass ? bot->sus : bot->su;
    const real_t* s2d = pass ? bot->sds : bot->sd;
    real_t* au = pass ? ans->sus : ans->su;
    real_t* ad = pass ? ans->sds : ans->sd;
    mv(n, bot->Rab, s1d, s->v1);
    for (int i = 0; i < n; ++i) au[i] = s1u[i] + t[i] * (s2u[i] + s->v1[i]);
    mv(n, bot->Tab, s1d, tmp);
    for (int i = 0; i < n; ++i) ad[i] = s2d[i] + tmp[i];
  }
  ans->scat = bot->scat;
}

/* src/adding.hpp:181-239 */
static void add_nonscat_bot(int n, const layer_t* top, const layer_t* bot, layer_t* ans,
                            addops_t* ops, scratch_t* s) {
  if (ops) ops->general = 0;
  real_t* t = s->v3;
  for (int i = 0; i < n; ++i) t[i] = bot->Tab[i * n + i];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      ans->Rab[i * n + j] = top->Rab[i * n + j];
      ans->Rba[i * n + j] = t[i] * top->Rba[i * n + j] * t[j];
      ans->Tab[i * n + j] = t[i] * top->Tab[i * n + j];
      ans->Tba[i * n + j] = top->Tba[i * n + j] * t[j];
    }
  real_t* tmp = s->t6;
  for (int pass = 0; pass < 2; ++pass) {
    const real_t* s1u = pass ? top->sus : top->su;
    const real_t* s1d = pass ? top->sds : top->sd;
    const real_t* s2u = pass ? bot->sus : bot->su;
    const real_t* s2d = pass ? bot->sds : bot->sd;
    real_t* au = pass ? ans->sus : ans->su;
    real_t* ad = pass ? ans->sds : ans->sd;
    mv(n, top->Tba, s2u, tmp);
    for (int i = 0; i < n; ++i) au[i] = s1u[i] + tmp[i];
    mv(n, top->Rba, s2u, s->v1);
    for (int i = 0; i < n; ++i) ad[i] = s2d[i] + t[i] * (s1d[i] + s->v1[i]);
  }
  ans->scat = top->scat;
}

/* src/adding.hpp:246-258 */
static void add_layers(int n, const layer_t* top, const layer_t* bot, layer_t* ans,
                       addops_t* ops, scratch_t* s) {
  memset(ans->su, 0, sizeof(real_t) * (size_t)(6 * n));
  if (!top->scat) add_nonscat_top(n, top, bot, ans, ops, s);
  else if (!bot->scat) add_nonscat_bot(n, top, bot, ans, ops, s);
  else add_general(n, top, bot, ans, ops, s);
}

/* ------------------------------------------------------------------------ */
/* interface intensities (src/solver.cpp:24-73)                              */
/* ------------------------------------------------------------------------ */

static void interface_up(int n, const layer_t* top, const layer_t* base, const real_t* Itd,
                         const real_t* Ibu, real_t* out, real_t* lu_keep, int* perm_keep,
                         scratch_t* s) {
  real_t* M = s->t1;
  mm(n, base->Rab, top->Rba, M);
  for (int e = 0; e < n * n; ++e) M[e] = -M[e];
  for (int i = 0; i < n; ++i) M[i * n + i] += 1.0;
  real_t* term1 = s->t2;        /* vectors held in matrix scratch */
  real_t* TtopI = s->t2 + n;
  real_t* term2 = s->t2 + 2 * n;
  real_t* sdt = s->t2 + 3 * n;
  real_t* term3 = s->t2 + 4 * n;
  real_t* rhs = s->t2 + 5 * n;
  mv(n, base->Tba, Ibu, term1);
  mv(n, top->Tab, Itd, TtopI);
  mv(n, base->Rab, TtopI, term2);
  for (int i = 0; i < n; ++i) sdt[i] = top->sd[i] + top->sds[i];
  mv(n, base->Rab, sdt, term3);
  for (int i = 0; i < n; ++i)
    rhs[i] = term1[i] + term2[i] + term3[i] + base->su[i] + base->sus[i];
  if (lu_keep) {
    memcpy(lu_keep, M, sizeof(real_t) * (size_t)(n * n));
    lu_factor(n, lu_keep, perm_keep);
    lu_solve(n, lu_keep, perm_keep, rhs, out);
  } else {
    int perm[ORC_MAX_N];
    solve_vec(n, M, rhs, out, s->work, perm);
  }
}

static void interface_down(int n, const layer_t* top, const layer_t* base,
                           const real_t* Itd, const real_t* Ibu, real_t* out,
                           real_t* lu_keep, int* perm_keep, scratch_t* s) {
  real_t* M = s->t1;
  mm(n, top->Rba, base->Rab, M);
  for (int e = 0; e < n * n; ++e) M[e] = -M[e];
  for (int i = 0; i < n; ++i) M[i * n + i] += 1.0;
  real_t* term1 = s->t2;
  real_t* TbI = s->t2 + n;
  real_t* term2 = s->t2 + 2 * n;
  real_t* sut = s->t2 + 3 * n;
  real_t* term3 = s->t2 + 4 * n;
  real_t* rhs = s->t2 + 5 * n;
  mv(n, top->Tab, Itd, term1);
  mv(n, base->Tba, Ibu, TbI);
  mv(n, top->Rba, TbI, term2);
  for (int i = 0; i < n; ++i) sut[i] = base->su[i] + base->sus[i];
  mv(n, top->Rba, sut, term3);
  for (int i = 0; i < n; ++i)
    rhs[i] = term1[i] + term2[i] + term3[i] + top->sd[i] + top->sds[i];
  if (lu_keep) {
    memcpy(lu_keep, M, sizeof(real_t) * (size_t)(n * n));
    lu_factor(n, lu_keep, perm_keep);
    lu_solve(n, lu_keep, perm_keep, rhs, out);
  } else {
    int perm[ORC_MAX_N];
    solve_vec(n, M, rhs, out, s->work, perm);
  }
}

/* ------------------------------------------------------------------------ */
/* validation (src/adding_doubling.hpp:123-174)                              */
/* ------------------------------------------------------------------------ */

static int validate(const orc_config* c, char* err, int errlen) {
  char msg[160];
  if (c->num_layers <= 0)
    return set_err(err, errlen, 1, "ADConfig: num_layers must be > 0");
  if (c->num_quadrature < 2)
    return set_err(err, errlen, 2, "ADConfig: num_quadrature must be >= 2");
  if (c->num_quadrature > ORC_MAX_N)
    return set_err(err, errlen, 17, "ADConfig: num_quadrature exceeds the supported maximum");
  if (!c->delta_tau) return set_err(err, errlen, 3, "ADConfig: delta_tau size mismatch");
  if (!c->single_scat_albedo)
    return set_err(err, errlen, 4, "ADConfig: single_scat_albedo size mismatch");
  if (!c->phase_moments || c->num_moments < 1)
    return set_err(err, errlen, 5, "ADConfig: phase_function_moments size mismatch");
  for (int l = 0; l < c->num_layers; ++l) {
    if (c->delta_tau[l] < 0.0) {
      snprintf(msg, sizeof msg, "ADConfig: delta_tau[%d] < 0", l);
      return set_err(err, errlen, 6, msg);
    }
    if (c->single_scat_albedo[l] < 0.0 || c->single_scat_albedo[l] > 1.0) {
      snprintf(msg, sizeof msg, "ADConfig: single_scat_albedo[%d] out of [0,1]", l);
      return set_err(err, errlen, 7, msg);
    }
  }
  if (c->surface_albedo < 0.0 || c->surface_albedo > 1.0)
    return set_err(err, errlen, 8, "ADConfig: surface_albedo out of [0,1]");
  if (c->solar_flux > 0.0 && (c->solar_mu <= 0.0 || c->solar_mu > 1.0))
    return set_err(err, errlen, 9, "ADConfig: solar_mu must be in (0,1] when solar_flux > 0");
  if (c->use_thermal_emission) {
    if (!c->temperature)
      return set_err(err, errlen, 10, "ADConfig: temperature size must be num_layers+1");
    for (int l = 0; l <= c->num_layers; ++l)
      if (c->temperature[l] < 0.0) {
        snprintf(msg, sizeof msg, "ADConfig: temperature[%d] < 0", l);
        return set_err(err, errlen, 11, msg);
      }
    if (c->wavenumber_low < 0.0 || c->wavenumber_high < c->wavenumber_low)
      return set_err(err, errlen, 12, "ADConfig: invalid wavenumber range");
  } else if (c->surface_temperature >= 0.0) {
    return set_err(err, errlen, 13,
                   "ADConfig: surface_temperature requires use_thermal_emission "
                   "(use surface_emission for raw thermal sources)");
  } else if (c->top_temperature >= 0.0) {
    return set_err(err, errlen, 14,
                   "ADConfig: top_temperature requires use_thermal_emission "
                   "(use top_emission for raw thermal sources)");
  }
  if (c->use_diffusion_lower_bc && !c->use_thermal_emission && !c->planck_levels)
    return set_err(err, errlen, 15,
                   "ADConfig: use_diffusion_lower_bc requires use_thermal_emission or "
                   "planck_levels");
  return 0;
}

/* ------------------------------------------------------------------------ */
/* column solve (src/solver.cpp:348-722)                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, nlay, ltot, has_surface;
  real_t *mu, *wt;
  real_t *B, *tau_used;
  layer_t *layers, *rbase, *rtop; /* layers[ltot], rbase/rtop[ltot+1] */
  real_t *mem_layers, *mem_rbase, *mem_rtop;
  addops_t *rbase_ops, *rtop_ops;
  real_t *iface_up_lu, *iface_dn_lu; /* [(nlay+1)][n*n] */
  int *iface_up_p, *iface_dn_p;      /* [(nlay+1)][ORC_MAX_N] */
  char* iface_has;
  real_t* ops_mem;
} column_t;

static void column_free(column_t* c) {
  free(c->mu);
  free(c->B);
  free(c->mem_layers);
  free(c->mem_rbase);
  free(c->mem_rtop);
  free(c->layers);
  free(c->rbase);
  free(c->rtop);
  free(c->rbase_ops);
  free(c->rtop_ops);
  free(c->iface_up_lu);
  free(c->iface_up_p);
  free(c->iface_has);
  free(c->ops_mem);
}

static void temperature_jacobian(const orc_config* cfg, column_t* C, const real_t* omega,
                                 const real_t* temp, int have_plev, orc_output* out);

int orc_solve(const orc_config* cfg_in, orc_output* out, char* err, int errlen) {
  int rc = validate(cfg_in, err, errlen);
  if (rc) return rc;
  const int n = cfg_in->num_quadrature;
  const int nlay = cfg_in->num_layers;
  const int M = cfg_in->num_moments;
  const int rev = cfg_in->index_from_bottom;
  const size_t nn2 = (size_t)(n * n);

  /* index_from_bottom reversal (src/solver.cpp:361-372) */
  real_t* tau = (real_t*)malloc(sizeof(real_t) * (size_t)nlay);
  real_t* omg = (real_t*)malloc(sizeof(real_t) * (size_t)nlay);
  real_t* mom = (real_t*)malloc(sizeof(real_t) * (size_t)(nlay * M));
  real_t* temp = NULL;
  real_t* plev = NULL;
  for (int l = 0; l < nlay; ++l) {
    int src = rev ? nlay - 1 - l : l;
    tau[l] = cfg_in->delta_tau[src];
    omg[l] = cfg_in->single_scat_albedo[src];
    for (int m = 0; m < M; ++m)
      mom[(size_t)l * M + m] = cfg_in->phase_moments[(size_t)src * M + m];
  }
  if (cfg_in->use_thermal_emission) {
    temp = (real_t*)malloc(sizeof(real_t) * (size_t)(nlay + 1));
    for (int l = 0; l <= nlay; ++l) temp[l] = cfg_in->temperature[rev ? nlay - l : l];
  }
  if (cfg_in->planck_levels) {
    plev = (real_t*)malloc(sizeof(real_t) * (size_t)(nlay + 1));
    for (int l = 0; l <= nlay; ++l) plev[l] = cfg_in->planck_levels[rev ? nlay - l : l];
  }
  /* scalars and flags from the caller; the (reversed) arrays are tau, omg,
   * mom, temp and plev above, in the internal arithmetic type */
  const orc_config* cfg = cfg_in;

  column_t C;
  memset(&C, 0, sizeof C);
  C.n = n;
  C.nlay = nlay;
  C.mu = (real_t*)malloc(sizeof(real_t) * (size_t)(2 * n));
  C.wt = C.mu + n;
  gl_r(n, C.mu, C.wt);
  const real_t* mu = C.mu;
  const real_t* wt = C.wt;
  real_t xfac_sum = 0.0;
  for (int i = 0; i < n; ++i) xfac_sum += mu[i] * wt[i];
  real_t xfac = 0.5 / xfac_sum;

  /* Planck values (src/solver.cpp:385-419) */
  C.B = (real_t*)calloc((size_t)(2 * nlay + 1), sizeof(real_t));
  C.tau_used = C.B + nlay + 1;
  real_t* B = C.B;
  real_t B_surface = 0.0, B_top_emission = 0.0;
  if (cfg->use_thermal_emission) {
    for (int l = 0; l <= nlay; ++l) {
      rc = planck_r(cfg->wavenumber_low, cfg->wavenumber_high, temp[l], &B[l]);
      if (rc) goto planck_fail;
    }
    if (cfg->surface_temperature >= 0.0) {
      rc = planck_r(cfg->wavenumber_low, cfg->wavenumber_high, cfg->surface_temperature,
                      &B_surface);
      if (rc) goto planck_fail;
    } else {
      B_surface = B[nlay];
    }
    if (cfg->top_temperature >= 0.0) {
      rc = planck_r(cfg->wavenumber_low, cfg->wavenumber_high, cfg->top_temperature,
                      &B_top_emission);
      if (rc) goto planck_fail;
    } else {
      B_top_emission = B[0];
    }
  } else if (plev) {
    for (int l = 0; l <= nlay; ++l) B[l] = plev[l];
    B_surface = cfg->surface_emission;
    B_top_emission = cfg->top_emission;
  } else {
    B_surface = cfg->surface_emission;
    B_top_emission = cfg->top_emission;
  }

  {
    /* per-layer R, T, sources (src/solver.cpp:421-514) */
    const int has_solar = (cfg->solar_flux > 0.0 && cfg->solar_mu > 0.0);
    const int twoM = 2 * n;
    int has_surface = !cfg->use_diffusion_lower_bc &&
                      (cfg->surface_albedo > 0.0 || B_surface > 0.0);
    int ltot = nlay + (has_surface ? 1 : 0);
    C.ltot = ltot;
    C.has_surface = has_surface;
    size_t ld = layer_doubles(n);
    C.mem_layers = (real_t*)malloc(sizeof(real_t) * ld * (size_t)ltot);
    C.mem_rbase = (real_t*)malloc(sizeof(real_t) * ld * (size_t)(ltot + 1));
    C.mem_rtop = (real_t*)malloc(sizeof(real_t) * ld * (size_t)(ltot + 1));
    C.layers = (layer_t*)malloc(sizeof(layer_t) * (size_t)ltot);
    C.rbase = (layer_t*)malloc(sizeof(layer_t) * (size_t)(ltot + 1));
    C.rtop = (layer_t*)malloc(sizeof(layer_t) * (size_t)(ltot + 1));
    for (int l = 0; l < ltot; ++l) layer_bind(&C.layers[l], C.mem_layers + ld * l, n);
    for (int l = 0; l <= ltot; ++l) {
      layer_bind(&C.rbase[l], C.mem_rbase + ld * l, n);
      layer_bind(&C.rtop[l], C.mem_rtop + ld * l, n);
      layer_default(&C.rbase[l], n);
      layer_default(&C.rtop[l], n);
    }

    real_t* Ppp = (real_t*)malloc(sizeof(real_t) * (nn2 * 2 + (size_t)(4 * n)));
    real_t* Ppm = Ppp + nn2;
    real_t* pps = Ppm + nn2;
    real_t* pms = pps + n;
    real_t* chis = (real_t*)malloc(sizeof(real_t) * (size_t)(M + twoM + 1));
    real_t* Pl = (real_t*)malloc(sizeof(real_t) * (size_t)((M + twoM + 1) * n));
    real_t tau_cumulative = 0.0;
    for (int l = 0; l < nlay; ++l) {
      real_t tau_layer = tau[l];
      real_t omega_layer = omg[l];
      const real_t* chi_full = mom + (size_t)l * M;
      memset(Ppp, 0, sizeof(real_t) * (nn2 * 2 + (size_t)(4 * n)));
      int has_solar_phase = 0;
      if (omega_layer > 0.0 && tau_layer > 0.0) {
        const real_t* chi = chi_full;
        int Lm = M;
        if (cfg->use_delta_m) {
          real_t f = (M > twoM) ? chi_full[twoM] : 0.0;
          if (f > 1e-12 && f < 1.0 - 1e-12) {
            real_t omega_f = omega_layer * f;
            tau_layer = (1.0 - omega_f) * tau[l];
            omega_layer = omega_layer * (1.0 - f) / (1.0 - omega_f);
            real_t inv_1mf = 1.0 / (1.0 - f);
            for (int ll = 0; ll < twoM; ++ll) chis[ll] = (chi_full[ll] - f) * inv_1mf;
            chi = chis;
            Lm = twoM;
          } else {
            Lm = (M < twoM) ? M : twoM;
          }
        }
        legtab_r(Lm, mu, n, Pl);
        phase_from_table(chi, Lm, Pl, n, wt, Ppp, Ppm);
        if (has_solar) {
          solar_from_table(chi, Lm, Pl, n, wt, cfg->solar_mu, pps, pms);
          has_solar_phase = 1;
        }
      }
      C.tau_used[l] = tau_layer;
      doubling_layer(n, tau_layer, omega_layer, B[l], B[l + 1], Ppp, Ppm, mu, wt,
                     cfg->solar_flux, cfg->solar_mu, tau_cumulative,
                     has_solar_phase ? pps : NULL, has_solar_phase ? pms : NULL,
                     &C.layers[l]);
      tau_cumulative += tau_layer;
    }
    free(Ppp);
    free(chis);
    free(Pl);

    /* Lambertian surface layer (src/solver.cpp:516-551) */
    if (has_surface) {
      layer_t* s = &C.layers[nlay];
      memset(s->Rab, 0, sizeof(real_t) * ld);
      real_t A = cfg->surface_albedo;
      for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
          real_t r = 2.0 * A * mu[j] * wt[j] * xfac;
          s->Rab[i * n + j] = r;
          s->Rba[i * n + j] = r;
        }
        s->su[i] = (1.0 - A) * B_surface;
        s->sd[i] = 0.0;
        if (has_solar && A > 0.0)
          s->sus[i] = (A / PI) * cfg->solar_flux * cfg->solar_mu *
                      exp(-tau_cumulative / cfg->solar_mu);
      }
      s->scat = (A > 0.0);
    }

    int want_jac = cfg->compute_temperature_jacobian &&
                   (cfg->use_thermal_emission || plev != NULL);
    if (want_jac) {
      C.rbase_ops = (addops_t*)calloc((size_t)ltot, sizeof(addops_t));
      C.rtop_ops = (addops_t*)calloc((size_t)ltot, sizeof(addops_t));
      C.ops_mem = (real_t*)malloc(sizeof(real_t) * nn2 * (size_t)(4 * ltot));
      for (int l = 0; l < ltot; ++l) {
        C.rbase_ops[l].lu1 = C.ops_mem + nn2 * (size_t)(4 * l);
        C.rbase_ops[l].lu2 = C.rbase_ops[l].lu1 + nn2;
        C.rtop_ops[l].lu1 = C.rbase_ops[l].lu2 + nn2;
        C.rtop_ops[l].lu2 = C.rtop_ops[l].lu1 + nn2;
      }
      C.iface_up_lu = (real_t*)malloc(sizeof(real_t) * nn2 * (size_t)(2 * (nlay + 1)));
      C.iface_dn_lu = C.iface_up_lu + nn2 * (size_t)(nlay + 1);
      C.iface_up_p = (int*)malloc(sizeof(int) * (size_t)(2 * (nlay + 1) * ORC_MAX_N));
      C.iface_dn_p = C.iface_up_p + (nlay + 1) * ORC_MAX_N;
      C.iface_has = (char*)calloc((size_t)(nlay + 1), 1);
    }

    scratch_t S;
    real_t* sblock;
    scratch_alloc(&S, n, &sblock);

    /* composites from the bottom (src/solver.cpp:560-572) */
    memcpy(C.rbase[1].Rab, C.layers[ltot - 1].Rab, sizeof(real_t) * ld);
    C.rbase[1].scat = C.layers[ltot - 1].scat;
    for (int l = 1; l < ltot; ++l) {
      int k = ltot - 1 - l;
      add_layers(n, &C.layers[k], &C.rbase[l], &C.rbase[l + 1],
                 want_jac ? &C.rbase_ops[l] : NULL, &S);
    }
    /* composites from the top (src/solver.cpp:574-583) */
    memcpy(C.rtop[1].Rab, C.layers[0].Rab, sizeof(real_t) * ld);
    C.rtop[1].scat = C.layers[0].scat;
    for (int l = 1; l < ltot; ++l)
      add_layers(n, &C.rtop[l], &C.layers[l], &C.rtop[l + 1],
                 want_jac ? &C.rtop_ops[l] : NULL, &S);

    /* boundary intensities (src/solver.cpp:585-603) */
    real_t Itd[ORC_MAX_N], Ibu[ORC_MAX_N];
    for (int i = 0; i < n; ++i) Itd[i] = B_top_emission;
    for (int i = 0; i < n; ++i) Ibu[i] = 0.0;
    if (cfg->use_diffusion_lower_bc) {
      real_t B_bottom = B[nlay];
      real_t dtau_last = C.tau_used[nlay - 1];
      real_t dB_dtau = (dtau_last > 0.0) ? (B_bottom - B[nlay - 1]) / dtau_last : 0.0;
      for (int i = 0; i < n; ++i) Ibu[i] = B_bottom + mu[i] * dB_dtau;
    } else if (!has_surface) {
      for (int i = 0; i < n; ++i) Ibu[i] = B_surface;
    }

    /* outputs (src/solver.cpp:605-700) */
    int nif = nlay + 1;
    real_t* res = (real_t*)calloc((size_t)(5 * nif), sizeof(real_t));
    real_t *fu = res, *fd = res + nif, *mi = res + 2 * nif, *fdiv = res + 3 * nif,
           *fdir = res + 4 * nif;
    if (cfg->solar_flux > 0.0 && cfg->solar_mu > 0.0) {
      real_t tau_cum = 0.0;
      fdir[0] = cfg->solar_flux * cfg->solar_mu;
      for (int l = 0; l < nlay; ++l) {
        tau_cum += C.tau_used[l];
        fdir[l + 1] = cfg->solar_flux * cfg->solar_mu * exp(-tau_cum / cfg->solar_mu);
      }
    }
    real_t Iup[ORC_MAX_N], Idn[ORC_MAX_N], tmpv[ORC_MAX_N];
    {
      layer_t* full = &C.rbase[ltot];
      mv(n, full->Rab, Itd, Iup);
      mv(n, full->Tba, Ibu, tmpv);
      for (int i = 0; i < n; ++i) Iup[i] += tmpv[i] + full->su[i] + full->sus[i];
      for (int i = 0; i < n; ++i) {
        fu[0] += 2.0 * PI * wt[i] * mu[i] * Iup[i];
        fd[0] += 2.0 * PI * wt[i] * mu[i] * Itd[i];
        mi[0] += 0.5 * wt[i] * (Iup[i] + Itd[i]);
      }
    }
    for (int l = 1; l <= nlay; ++l) {
      int n_top = l, n_base = ltot - l;
      if (n_base > 0 && n_top > 0) {
        interface_up(n, &C.rtop[n_top], &C.rbase[n_base], Itd, Ibu, Iup,
                     want_jac ? C.iface_up_lu + nn2 * l : NULL,
                     want_jac ? C.iface_up_p + l * ORC_MAX_N : NULL, &S);
        interface_down(n, &C.rtop[n_top], &C.rbase[n_base], Itd, Ibu, Idn,
                       want_jac ? C.iface_dn_lu + nn2 * l : NULL,
                       want_jac ? C.iface_dn_p + l * ORC_MAX_N : NULL, &S);
        if (want_jac) C.iface_has[l] = 1;
      } else if (n_base == 0) {
        layer_t* top = &C.rtop[n_top];
        mv(n, top->Tab, Itd, Idn);
        mv(n, top->Rba, Ibu, tmpv);
        for (int i = 0; i < n; ++i) Idn[i] += tmpv[i] + top->sd[i] + top->sds[i];
        for (int i = 0; i < n; ++i) Iup[i] = Ibu[i];
      } else {
        for (int i = 0; i < n; ++i) { Iup[i] = 0.0; Idn[i] = Itd[i]; }
      }
      for (int i = 0; i < n; ++i) {
        fu[l] += 2.0 * PI * wt[i] * mu[i] * Iup[i];
        fd[l] += 2.0 * PI * wt[i] * mu[i] * Idn[i];
        mi[l] += 0.5 * wt[i] * (Iup[i] + Idn[i]);
      }
    }
    if (cfg->solar_flux > 0.0 && cfg->solar_mu > 0.0) {
      real_t inv_4pi_mu0 = 1.0 / (4.0 * PI * cfg->solar_mu);
      for (int l = 0; l < nif; ++l) mi[l] += fdir[l] * inv_4pi_mu0;
    }
    for (int l = 0; l < nif; ++l) {
      int lyr = (l == 0) ? 0 : l - 1;
      fdiv[l] = 4.0 * PI * (1.0 - omg[lyr]) * (mi[l] - B[l]);
    }

    if (want_jac) temperature_jacobian(cfg, &C, omg, temp, plev != NULL, out);

    double* dst[5] = {out->flux_up, out->flux_down, out->mean_intensity,
                      out->flux_divergence, out->flux_direct};
    for (int a = 0; a < 5; ++a)
      if (dst[a])
        for (int l = 0; l < nif; ++l) dst[a][l] = (double)res[a * nif + (rev ? nif - 1 - l : l)];
    free(res);
    free(sblock);
  }
  column_free(&C);
  free(tau);
  free(omg);
  free(mom);
  free(temp);
  free(plev);
  return 0;

planck_fail:
  column_free(&C);
  free(tau);
  free(omg);
  free(mom);
  free(temp);
  free(plev);
  return set_err(err, errlen, rc, "planckFunction: invalid arguments");
}

/* ------------------------------------------------------------------------ */
/* temperature Jacobian (src/solver.cpp:117-341)                             */
/* ------------------------------------------------------------------------ */

/* blocks are [n][ndof] row-major */
static void temperature_jacobian(const orc_config* cfg, column_t* C, const real_t* omega,
                                 const real_t* temp, int have_plev, orc_output* out) {
  const int n = C->n, nlay = C->nlay, ltot = C->ltot, has_surface = C->has_surface;
  const int thermal = cfg->use_thermal_emission;
  const int nif = nlay + 1, ndof = nlay + 2, surf_dof = nlay + 1;
  const real_t A = cfg->surface_albedo;
  const real_t *mu = C->mu, *wt = C->wt;
  const int surface_independent =
      has_surface && (thermal ? (cfg->surface_temperature >= 0.0) : have_plev);
  const int surf_col = surface_independent ? surf_dof : nlay;
  const size_t bs = (size_t)(n * ndof);

  real_t* mem = (real_t*)calloc(bs * (size_t)(4 * (ltot + 1) + 12), sizeof(real_t));
  real_t* dRbu = mem;
  real_t* dRbd = dRbu + bs * (ltot + 1);
  real_t* dRtu = dRbd + bs * (ltot + 1);
  real_t* dRtd = dRtu + bs * (ltot + 1);
  real_t* su = dRtd + bs * (ltot + 1);
  real_t* sd = su + bs;
  real_t* tU = sd + bs;
  real_t* tD = tU + bs;
  real_t* dItd = tD + bs;
  real_t* dIbu = dItd + bs;
  real_t* dIup = dIbu + bs;
  real_t* dIdn = dIup + bs;
  real_t* tmp1 = dIdn + bs;
  real_t* tmp2 = tmp1 + bs;
  real_t* colv = tmp2 + bs;       /* n */
  real_t* colx = colv + n;        /* n */
  (void)colx;

#define BLK(P, l) ((P) + bs * (size_t)(l))

  /* seedLayer (src/solver.cpp:150-165) */
#define SEED(j, dU, dD)                                                   \
  do {                                                                    \
    memset(dU, 0, sizeof(real_t) * bs);                                   \
    memset(dD, 0, sizeof(real_t) * bs);                                   \
    if (has_surface && (j) == nlay) {                                     \
      for (int i = 0; i < n; ++i) (dU)[i * ndof + surf_col] = (1.0 - A);  \
    } else {                                                              \
      const layer_t* Ly = &C->layers[(j)];                                \
      for (int i = 0; i < n; ++i) {                                       \
        (dU)[i * ndof + (j)] += Ly->jp[i];                                \
        (dU)[i * ndof + (j) + 1] += Ly->jq[i];                            \
        (dD)[i * ndof + (j)] += Ly->jq[i];                                \
        (dD)[i * ndof + (j) + 1] += Ly->jp[i];                            \
      }                                                                   \
    }                                                                     \
  } while (0)

  /* Y = M X for M [n][n], X [n][ndof] */
#define MB(Mat, X, Y)                                                     \
  do {                                                                    \
    const real_t* mb_a_ = (Mat);                                          \
    const real_t* mb_x_ = (X);                                            \
    real_t* mb_y_ = (Y);                                                  \
    for (int mb_i_ = 0; mb_i_ < n; ++mb_i_)                               \
      for (int mb_m_ = 0; mb_m_ < ndof; ++mb_m_) {                        \
        real_t mb_s_ = 0.0;                                               \
        for (int mb_k_ = 0; mb_k_ < n; ++mb_k_)                           \
          mb_s_ += mb_a_[mb_i_ * n + mb_k_] * mb_x_[mb_k_ * ndof + mb_m_]; \
        mb_y_[mb_i_ * ndof + mb_m_] = mb_s_;                              \
      }                                                                   \
  } while (0)

  /* columnwise LU solve of a block in place */
#define LUB(LU, P, X)                                                     \
  do {                                                                    \
    for (int m = 0; m < ndof; ++m) {                                      \
      for (int i = 0; i < n; ++i) colv[i] = (X)[i * ndof + m];            \
      real_t xs[ORC_MAX_N];                                               \
      lu_solve(n, (LU), (P), colv, xs);                                   \
      for (int i = 0; i < n; ++i) (X)[i * ndof + m] = xs[i];              \
    }                                                                     \
  } while (0)

  /* addSourcesLin (src/solver.cpp:173-190) */
#define ADDLIN(top, bot, ops, dUt, dDt, dUb, dDb, dUa, dDa)               \
  do {                                                                    \
    MB((bot)->Rab, dDt, tmp1);                                            \
    for (size_t e = 0; e < bs; ++e) tU[e] = (dUb)[e] + tmp1[e];          \
    MB((top)->Rba, dUb, tmp1);                                            \
    for (size_t e = 0; e < bs; ++e) tD[e] = (dDt)[e] + tmp1[e];          \
    if ((ops)->general) {                                                 \
      LUB((ops)->lu1, (ops)->p1, tU);                                     \
      LUB((ops)->lu2, (ops)->p2, tD);                                     \
    }                                                                     \
    MB((top)->Tba, tU, tmp1);                                             \
    MB((bot)->Tab, tD, tmp2);                                             \
    for (size_t e = 0; e < bs; ++e) {                                     \
      (dUa)[e] = (dUt)[e] + tmp1[e];                                      \
      (dDa)[e] = (dDb)[e] + tmp2[e];                                      \
    }                                                                     \
  } while (0)

  SEED(ltot - 1, BLK(dRbu, 1), BLK(dRbd, 1));
  for (int l = 1; l < ltot; ++l) {
    int k = ltot - 1 - l;
    SEED(k, su, sd);
    ADDLIN(&C->layers[k], &C->rbase[l], &C->rbase_ops[l], su, sd, BLK(dRbu, l),
           BLK(dRbd, l), BLK(dRbu, l + 1), BLK(dRbd, l + 1));
  }
  SEED(0, BLK(dRtu, 1), BLK(dRtd, 1));
  for (int l = 1; l < ltot; ++l) {
    SEED(l, su, sd);
    ADDLIN(&C->rtop[l], &C->layers[l], &C->rtop_ops[l], BLK(dRtu, l), BLK(dRtd, l), su, sd,
           BLK(dRtu, l + 1), BLK(dRtd, l + 1));
  }

  /* boundary-intensity derivatives (src/solver.cpp:214-235) */
  if (thermal && cfg->top_temperature < 0.0)
    for (int i = 0; i < n; ++i) dItd[i * ndof + 0] = 1.0;
  if (!has_surface && cfg->use_diffusion_lower_bc) {
    real_t dtau_last = C->tau_used[nlay - 1];
    if (dtau_last > 0.0) {
      for (int i = 0; i < n; ++i) {
        dIbu[i * ndof + nlay] = 1.0 + mu[i] / dtau_last;
        dIbu[i * ndof + nlay - 1] = -mu[i] / dtau_last;
      }
    } else {
      for (int i = 0; i < n; ++i) dIbu[i * ndof + nlay] = 1.0;
    }
  }

  size_t jsz = (size_t)(nif * ndof);
  real_t* J = (real_t*)calloc(jsz * 4, sizeof(real_t));
  real_t *Ju = J, *Jd = J + jsz, *Jm = J + 2 * jsz, *Jdv = J + 3 * jsz;

#define REDUCE(k, U, D)                                                   \
  do {                                                                    \
    for (int m = 0; m < ndof; ++m) {                                      \
      real_t fu = 0.0, fd = 0.0, jm = 0.0;                                \
      for (int i = 0; i < n; ++i) {                                       \
        fu += 2.0 * PI * wt[i] * mu[i] * (U)[i * ndof + m];               \
        fd += 2.0 * PI * wt[i] * mu[i] * (D)[i * ndof + m];               \
        jm += 0.5 * wt[i] * ((U)[i * ndof + m] + (D)[i * ndof + m]);      \
      }                                                                   \
      Ju[(k) * ndof + m] = fu;                                            \
      Jd[(k) * ndof + m] = fd;                                            \
      Jm[(k) * ndof + m] = jm;                                            \
    }                                                                     \
  } while (0)

  {
    layer_t* full = &C->rbase[ltot];
    MB(full->Rab, dItd, dIup);
    MB(full->Tba, dIbu, tmp1);
    real_t* dr = BLK(dRbu, ltot);
    for (size_t e = 0; e < bs; ++e) dIup[e] = dIup[e] + tmp1[e] + dr[e];
    REDUCE(0, dIup, dItd);
  }
  for (int l = 1; l <= nlay; ++l) {
    int n_top = l, n_base = ltot - l;
    if (n_base > 0) {
      layer_t* top = &C->rtop[n_top];
      layer_t* base = &C->rbase[n_base];
      MB(base->Tba, dIbu, dIup);
      MB(top->Tab, dItd, tmp1);
      MB(base->Rab, tmp1, tmp2);
      for (size_t e = 0; e < bs; ++e) dIup[e] += tmp2[e];
      MB(base->Rab, BLK(dRtd, n_top), tmp2);
      real_t* drb = BLK(dRbu, n_base);
      for (size_t e = 0; e < bs; ++e) dIup[e] += tmp2[e] + drb[e];
      LUB(C->iface_up_lu + (size_t)(n * n) * l, C->iface_up_p + l * ORC_MAX_N, dIup);

      MB(top->Tab, dItd, dIdn);
      MB(base->Tba, dIbu, tmp1);
      MB(top->Rba, tmp1, tmp2);
      for (size_t e = 0; e < bs; ++e) dIdn[e] += tmp2[e];
      MB(top->Rba, BLK(dRbu, n_base), tmp2);
      real_t* drt = BLK(dRtd, n_top);
      for (size_t e = 0; e < bs; ++e) dIdn[e] += tmp2[e] + drt[e];
      LUB(C->iface_dn_lu + (size_t)(n * n) * l, C->iface_dn_p + l * ORC_MAX_N, dIdn);
    } else {
      layer_t* top = &C->rtop[n_top];
      MB(top->Tab, dItd, dIdn);
      MB(top->Rba, dIbu, tmp1);
      real_t* drt = BLK(dRtd, n_top);
      for (size_t e = 0; e < bs; ++e) dIdn[e] += tmp1[e] + drt[e];
      memcpy(dIup, dIbu, sizeof(real_t) * bs);
    }
    REDUCE(l, dIup, dIdn);
  }

  /* flux-divergence Jacobian (src/solver.cpp:297-305) */
  for (int k = 0; k < nif; ++k) {
    int lyr = (k == 0) ? 0 : k - 1;
    real_t pref = 4.0 * PI * (1.0 - omega[lyr]);
    for (int m = 0; m < ndof; ++m)
      Jdv[k * ndof + m] = pref * (Jm[k * ndof + m] - ((m == k) ? 1.0 : 0.0));
  }

  /* Planck chain rule (src/solver.cpp:307-327) */
  real_t* chain = (real_t*)malloc(sizeof(real_t) * (size_t)ndof);
  for (int m = 0; m < ndof; ++m) chain[m] = 1.0;
  if (thermal) {
    for (int m = 0; m <= nlay; ++m)
      planckd_r(cfg->wavenumber_low, cfg->wavenumber_high, temp[m],
                       &chain[m]);
    if (surface_independent)
      planckd_r(cfg->wavenumber_low, cfg->wavenumber_high, cfg->surface_temperature,
                       &chain[surf_dof]);
    else
      chain[surf_dof] = 0.0;
  } else if (!surface_independent) {
    chain[surf_dof] = 0.0;
  }
  for (int a = 0; a < 4; ++a)
    for (int k = 0; k < nif; ++k)
      for (int m = 0; m < ndof; ++m) J[a * jsz + k * ndof + m] *= chain[m];

  /* index_from_bottom: reverse interface rows and level columns
   * (src/solver.cpp:329-340) */
  double* dst[4] = {out->jac_flux_up, out->jac_flux_down, out->jac_mean_intensity,
                    out->jac_flux_divergence};
  int rev = cfg->index_from_bottom;
  for (int a = 0; a < 4; ++a) {
    if (!dst[a]) continue;
    for (int k = 0; k < nif; ++k) {
      int ks = rev ? nif - 1 - k : k;
      for (int m = 0; m < ndof; ++m) {
        int ms = (rev && m <= nlay) ? nlay - m : m;
        dst[a][k * ndof + m] = (double)J[a * jsz + ks * ndof + ms];
      }
    }
  }
  free(chain);
  free(J);
  free(mem);
#undef BLK
#undef SEED
#undef MB
#undef LUB
#undef ADDLIN
#undef REDUCE
}

/* ------------------------------------------------------------------------ */
/* batched driver (CPU baseline and parity harness)                          */
/* ------------------------------------------------------------------------ */

typedef struct {
  const orc_batch* b;
  int w0, w1;
  int rc;
  char err[160];
} batch_job_t;

static void* batch_worker(void* arg) {
  batch_job_t* job = (batch_job_t*)arg;
  const orc_batch* b = job->b;
  const int L = b->num_layers, M = b->num_moments, nif = L + 1;
  orc_config c;
  memset(&c, 0, sizeof c);
  c.num_layers = L;
  c.num_quadrature = b->num_quadrature;
  c.num_moments = M;
  c.use_delta_m = b->use_delta_m;
  c.use_diffusion_lower_bc = b->use_diffusion_lower_bc;
  c.index_from_bottom = b->index_from_bottom;
  c.surface_albedo = b->surface_albedo;
  c.surface_temperature = -1.0;
  c.top_temperature = -1.0;
  c.solar_flux = b->solar_flux;
  c.solar_mu = b->solar_mu;
  double* zeros = NULL;
  if (!b->planck_levels && b->use_diffusion_lower_bc) {
    zeros = (double*)calloc((size_t)nif, sizeof(double));
  }
  for (int w = job->w0; w < job->w1; ++w) {
    c.delta_tau = b->delta_tau + (size_t)w * L;
    c.single_scat_albedo = b->single_scat_albedo + (size_t)w * L;
    c.phase_moments = b->phase_moments_shared ? b->phase_moments
                                              : b->phase_moments + (size_t)w * L * M;
    c.planck_levels = b->planck_levels ? b->planck_levels + (size_t)w * nif : zeros;
    c.surface_emission = b->surface_emission_w ? b->surface_emission_w[w] : b->surface_emission;
    c.top_emission = b->top_emission_w ? b->top_emission_w[w] : b->top_emission;
    orc_output o;
    memset(&o, 0, sizeof o);
    o.flux_up = b->flux_up + (size_t)w * nif;
    o.flux_down = b->flux_down + (size_t)w * nif;
    o.mean_intensity = b->mean_intensity + (size_t)w * nif;
    o.flux_direct = b->flux_direct + (size_t)w * nif;
    o.flux_divergence = b->flux_divergence ? b->flux_divergence + (size_t)w * nif : NULL;
    int rc = orc_solve(&c, &o, job->err, (int)sizeof job->err);
    if (rc) { job->rc = rc; break; }
    if (b->spectrum_scaling != 1.0)
      for (int l = 0; l < nif; ++l) o.flux_up[l] *= b->spectrum_scaling;
  }
  free(zeros);
  return NULL;
}

int orc_solve_batch(const orc_batch* b, int num_threads, char* err, int errlen) {
  int W = b->num_wavenumbers;
  if (W <= 0) return 0;
  if (num_threads < 1) num_threads = 1;
  if (num_threads > W) num_threads = W;
  batch_job_t* jobs = (batch_job_t*)calloc((size_t)num_threads, sizeof(batch_job_t));
  pthread_t* th = (pthread_t*)calloc((size_t)num_threads, sizeof(pthread_t));
  for (int t = 0; t < num_threads; ++t) {
    jobs[t].b = b;
    jobs[t].w0 = (int)((long long)W * t / num_threads);
    jobs[t].w1 = (int)((long long)W * (t + 1) / num_threads);
  }
  for (int t = 1; t < num_threads; ++t) pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
  batch_worker(&jobs[0]);
  for (int t = 1; t < num_threads; ++t) pthread_join(th[t], NULL);
  int rc = 0;
  for (int t = 0; t < num_threads && !rc; ++t)
    if (jobs[t].rc) rc = set_err(err, errlen, jobs[t].rc, jobs[t].err);
  free(jobs);
  free(th);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* exported building blocks (double interface)                               */
/* ------------------------------------------------------------------------ */

static real_t* to_r(const double* x, int n) {
  if (!x) return NULL;
  real_t* r = (real_t*)malloc(sizeof(real_t) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) r[i] = x[i];
  return r;
}

static void from_r(double* dst, const real_t* src, int n) {
  for (int i = 0; i < n; ++i) dst[i] = (double)src[i];
}

void orc_gauss_legendre(int n, double* nodes, double* weights) {
  real_t* a = to_r(nodes, n);
  real_t* b = to_r(weights, n);
  gl_r(n, a, b);
  from_r(nodes, a, n);
  from_r(weights, b, n);
  free(a);
  free(b);
}

void orc_legendre_table(int L, const double* x, int nx, double* Pl) {
  real_t* xr = to_r(x, nx);
  real_t* P = (real_t*)malloc(sizeof(real_t) * (size_t)(L * nx + 1));
  legtab_r(L, xr, nx, P);
  from_r(Pl, P, L * nx);
  free(xr);
  free(P);
}

void orc_phase_matrices(const double* chi, int L, const double* mu, const double* wt, int n,
                        double* Ppp, double* Ppm) {
  real_t *c = to_r(chi, L), *m = to_r(mu, n), *w = to_r(wt, n);
  real_t* A = (real_t*)malloc(sizeof(real_t) * (size_t)(2 * n * n));
  phase_r(c, L, m, w, n, A, A + n * n);
  from_r(Ppp, A, n * n);
  from_r(Ppm, A + n * n, n * n);
  free(c); free(m); free(w); free(A);
}

void orc_solar_phase_vectors(const double* chi, int L, const double* mu, const double* wt,
                             int n, double mu0, double* p_plus, double* p_minus) {
  real_t *c = to_r(chi, L), *m = to_r(mu, n), *w = to_r(wt, n);
  real_t* v = (real_t*)malloc(sizeof(real_t) * (size_t)(2 * n));
  solar_r(c, L, m, w, n, mu0, v, v + n);
  from_r(p_plus, v, n);
  from_r(p_minus, v + n, n);
  free(c); free(m); free(w); free(v);
}

int orc_planck(double lo, double hi, double T, double* out) {
  real_t r = 0;
  int rc = planck_r(lo, hi, T, &r);
  *out = (double)r;
  return rc;
}

int orc_planck_deriv(double lo, double hi, double T, double* out) {
  real_t r = 0;
  int rc = planckd_r(lo, hi, T, &r);
  *out = (double)r;
  return rc;
}

int orc_doubling_count(double tau, double omega) { return nn_r(tau, omega); }

int orc_doubling(int n, double tau, double omega, double B_top, double B_bottom,
                 const double* Ppp, const double* Ppm, const double* mu, const double* wt,
                 double solar_flux, double solar_mu, double tau_cumulative,
                 const double* p_plus_solar, const double* p_minus_solar, double* R, double* T,
                 double* s_up, double* s_down, double* s_up_solar, double* s_down_solar,
                 double* j_p, double* j_q) {
  real_t *pp = to_r(Ppp, n * n), *pm = to_r(Ppm, n * n), *m = to_r(mu, n), *w = to_r(wt, n);
  real_t *ps = to_r(p_plus_solar, n), *ms = to_r(p_minus_solar, n);
  real_t* o = (real_t*)malloc(sizeof(real_t) * (size_t)(2 * n * n + 6 * n));
  int scat = doubling_r(n, tau, omega, B_top, B_bottom, pp, pm, m, w, solar_flux, solar_mu,
                        tau_cumulative, ps, ms, o, o + n * n, o + 2 * n * n, o + 2 * n * n + n,
                        o + 2 * n * n + 2 * n, o + 2 * n * n + 3 * n, o + 2 * n * n + 4 * n,
                        o + 2 * n * n + 5 * n);
  from_r(R, o, n * n);
  from_r(T, o + n * n, n * n);
  double* vs[6] = {s_up, s_down, s_up_solar, s_down_solar, j_p, j_q};
  for (int k = 0; k < 6; ++k) from_r(vs[k], o + 2 * n * n + k * n, n);
  free(pp); free(pm); free(m); free(w); free(ps); free(ms); free(o);
  return scat;
}
