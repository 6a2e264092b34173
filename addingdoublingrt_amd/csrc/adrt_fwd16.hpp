// adrt_fwd16.hpp -- forward kernels for N = 16 on the matrix cores (gfx950).
//
// Same pipeline and arithmetic as adrt_fwd.hpp (the column-per-lane VALU
// kernels used for the other quadrature orders), laid out one problem per
// wave with the doubling / adding products on v_mfma_f64_16x16x4_f64.
// Records in the layer store hold R and T in register (lane) order.
#pragma once

#include "adrt_fwd.hpp"
#include "adrt_wave16.hpp"

namespace adrt {
namespace dev {
namespace w16 {

// Per-wave LDS region of the N = 16 layer and adding kernels (a workgroup of
// four waves holds four consecutive regions).  Every path -- Newton-Schulz,
// the LU fallback (which runs in registers and lane permutes, no LDS), the
// interface solves -- stays inside its wave's region: a write past it would
// land in the next wave's tiles, which that wave reads without any fault.
struct Lds16 {
  static constexpr int tile = 0, tile2 = TILE, pv = 2 * TILE, vs = pv + N;  // 4 vector slots
  static constexpr int size = vs + 4 * N;
};
static_assert(Lds16::tile + TILE <= Lds16::tile2 && Lds16::tile2 + TILE <= Lds16::pv &&
                  Lds16::pv + N <= Lds16::vs && Lds16::vs + 4 * N <= Lds16::size,
              "N = 16 per-wave LDS regions overlap");

constexpr int REC16 = 2 * N * N + 2 * N;  // R, T (lane order), s_up, s_down
// per layer, for the downward sweep of the adding: G_l (lane order; the
// diagonal of T_l for a non-scattering layer), the base below it Rb(l+1)
// (lane order), d_l, sb(l+1), and whether Rb(l+1) is non-zero
constexpr int RB_G = 0, RB_R = N * N, RB_D = 2 * N * N, RB_S = RB_D + N, RB_F = RB_S + N;
constexpr int RBR16 = 2 * N * N + 4 * N;
static_assert(RB_F < RBR16, "adding record fields past the record");

// ---------------------------------------------------------------------------
// phase matrices and thin-layer start of a scattering layer, in FP64, for the
// register layout of R (src/phase_matrix.hpp:61-181, src/doubling.hpp:79-147)
// ---------------------------------------------------------------------------
template <typename R>
struct Start16 {
  double Tv[4], Rv[4];  // T0, R0 at rows row<R>(h, r) of column j
  double Sv[4];         // I - T0 = tau0 G++, without cancellation
  double y, sus, sds;   // element j
  double gam, gk, tau0;
  int nn;
};

// Phase sums of one layer, shared by every wavenumber when the moments are
// (src/phase_matrix.hpp:61-181 before the Hansen normalisation): Qpp, Qpm
// (row-major, divided by 2 pi), the column sums csum[j] = sum_i (Qpp + Qpm)(i, j)
// w_i, the solar vectors q+, q- (divided by 2 pi) and ssol = sum_j (q+ + q-)_j w_j.
constexpr int PH16 = 2 * N * N + 4 * N;
constexpr int PH_QPM = N * N, PH_CSUM = 2 * N * N, PH_QP = 2 * N * N + N,
              PH_QM = 2 * N * N + 2 * N, PH_SSOL = 2 * N * N + 3 * N;

// Legendre sums of layer l's moments for the rows row<R>(h, r) of column j;
// returns the number of moments used (Lm) and the delta-M fraction f.
template <typename R>
__device__ __forceinline__ int phase16_sums(const FwdArgs& a, const double* chi, int mode,
                                            double f, const Lane& L, double (&Qpp)[4],
                                            double (&Qpm)[4], double& spj, double& smj) {
  const int j = L.j;
  const int twoM = 2 * a.n_real;
  int Lm = a.M;
  double inv1mf = 1.0;
  if (mode == 3) {
    inv1mf = 1.0 / (1.0 - f);
    Lm = twoM;
  } else if (mode == 2) {
    Lm = a.M < twoM ? a.M : twoM;
  }
  double sign = 1.0;
  spj = smj = 0.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) Qpp[r] = Qpm[r] = 0.0;
  for (int ll = 0; ll < Lm; ++ll) {
    const double chil = (mode == 3) ? (chi[ll] - f) * inv1mf : chi[ll];
    const double cl = (2 * ll + 1) * chil;
    const double* P = a.Pl + ll * N;
    const double plj = P[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double term = cl * P[row<R>(L.h, r)] * plj;
      Qpp[r] += term;
      Qpm[r] += sign * term;
    }
    if (a.has_solar) {
      const double term = cl * plj * a.Pl0[ll];
      spj += term;
      smj += sign * term;
    }
    sign = -sign;
  }
  return Lm;
}

// One wave per layer: the phase record of layer l (shared moments only).  The
// delta-M mode of a scattering layer depends on its moments alone
// (src/solver.cpp:445-486), so it is decided here as the prep kernel does.
__global__ void __launch_bounds__(64) phase16_kernel(FwdArgs a, double* phase) {
  __shared__ double tile[TILE];
  const int l = blockIdx.x;
  if (l >= a.L) return;
  const Lane L;
  const int j = L.j;
  const int li = lin(a, l);
  const double* chi = a.mom + (long)li * a.M;
  const int twoM = 2 * a.n_real;
  int mode = 1;
  double f = 0.0;
  if (a.delta_m) {
    f = (a.M > twoM) ? chi[twoM] : 0.0;
    mode = (f > 1e-12 && f < 1.0 - 1e-12) ? 3 : 2;
  }
  double Qpp[4], Qpm[4], spj, smj;
  phase16_sums<double>(a, chi, mode, f, L, Qpp, Qpm, spj, smj);
  double* ph = phase + (long)l * PH16;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = row<double>(L.h, r);
    Qpp[r] = Qpp[r] / (2.0 * kPi);
    Qpm[r] = Qpm[r] / (2.0 * kPi);
    ph[i * N + j] = Qpp[r];
    ph[PH_QPM + i * N + j] = Qpm[r];
    tile[j * TS + i] = (Qpp[r] + Qpm[r]) * a.wt[i];
  }
  wave_sync();
  double csum = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) csum += tile[j * TS + i];
  if (L.h == 0) {
    ph[PH_CSUM + j] = csum;
    ph[PH_QP + j] = spj / (2.0 * kPi);
    ph[PH_QM + j] = smj / (2.0 * kPi);
  }
  wave_sync();
  const double ssol = colsum<double>((spj / (2.0 * kPi) + smj / (2.0 * kPi)) * a.wt[j], L, tile);
  if (L.l == 0) ph[PH_SSOL] = ssol;
}

// dtile: a TILE of doubles (LDS), pv: 16 doubles (LDS)
template <typename R>
__device__ __forceinline__ void layer16_start(const FwdArgs& a, long task, int w, int l, long gw,
                                              const Lane& L, double tau, double omega_in,
                                              int mode, double* dtile, double* pv,
                                              Start16<R>& st) {
  const int L_ = a.L;
  const int j = L.j;
  const double mu_j = a.mu[j], wt_j = a.wt[j];
  const double omega = omega_in > 1.0 ? 1.0 : omega_in;
  const double con = 2.0 * omega * kPi;
  const int li = lin(a, l);
  const double* chi = a.shared_moments ? a.mom + (long)li * a.M : a.mom + (gw * L_ + li) * a.M;
  const double f = (mode == 3) ? a.fdm[task] : 0.0;
  double Ppp[4], Ppm[4];
  int rows[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) rows[r] = row<R>(L.h, r);
  double pplus = 0.0, pminus = 0.0;
  if (a.phase) {
    // shared moments: the layer's phase record (phase16_kernel)
    const double* ph = a.phase + (long)l * PH16;
    const double csum = ph[PH_CSUM + j];
    const double corr = csum > 0.0 ? 1.0 / (2.0 * kPi * csum) : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Ppp[r] = ph[rows[r] * N + j] * corr;
      Ppm[r] = ph[PH_QPM + rows[r] * N + j] * corr;
    }
    if (a.has_solar) {
      pplus = ph[PH_QP + j];
      pminus = ph[PH_QM + j];
      const double s = ph[PH_SSOL];
      if (s > 0.0) {
        const double c = 1.0 / (2.0 * kPi * s);
        pplus *= c;
        pminus *= c;
      }
    }
  } else {
    double spj, smj;
    phase16_sums<R>(a, chi, mode, f, L, Ppp, Ppm, spj, smj);
    // Hansen column normalisation, rows summed in order 0..15
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Ppp[r] = Ppp[r] / (2.0 * kPi);
      Ppm[r] = Ppm[r] / (2.0 * kPi);
      dtile[j * TS + rows[r]] = (Ppp[r] + Ppm[r]) * a.wt[rows[r]];
    }
    wave_sync();
    double csum = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) csum += dtile[j * TS + i];
    wave_sync();
    if (csum > 0.0) {
      const double corr = 1.0 / (2.0 * kPi * csum);
#pragma unroll
      for (int r = 0; r < 4; ++r) { Ppp[r] *= corr; Ppm[r] *= corr; }
    }
    if (a.has_solar) {
      pplus = spj / (2.0 * kPi);
      pminus = smj / (2.0 * kPi);
      const double s = colsum<double>((pplus + pminus) * wt_j, L, pv);
      if (s > 0.0) {
        const double corr = 1.0 / (2.0 * kPi * s);
        pplus *= corr;
        pminus *= corr;
      }
    }
  }
  // thin-layer initialisation (src/doubling.hpp:83-147)
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(omega);
  st.nn = nn0 < 1 ? 1 : nn0;
  st.tau0 = tau * ldexp(1.0, -st.nn);
  const double tau0 = st.tau0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = rows[r];
    const double mui = a.mu[i];
    const double gpp = (((i == j) ? 1.0 : 0.0) + (-con) * (Ppp[r] * wt_j)) / mui;
    const double gpm = con * (Ppm[r] * wt_j) / mui;
    st.Tv[r] = ((i == j) ? 1.0 : 0.0) - tau0 * gpp;
    st.Sv[r] = tau0 * gpp;
    st.Rv[r] = tau0 * gpm;
  }
  st.y = (1.0 - omega) * tau0 / mu_j;
  st.sus = 0.0;
  st.sds = 0.0;
  st.gam = 0.0;
  if (a.has_solar) {
    const double Ftop = a.F0 * exp(-a.tau_cum[(long)w * (L_ + 1) + l] / a.mu0);
    const double base = omega * tau0 / mu_j * Ftop;
    st.sus = base * pminus;
    st.sds = base * pplus;
    st.gam = exp(-tau0 / a.mu0);
  }
  st.gk = 0.5 * tau0;
}

// ---------------------------------------------------------------------------
// layer setup + doubling, one wave per (wavenumber, layer).
// REDO: only the layers the mixed-precision kernel marked (its FP32 recursion
// cannot follow them, see adrt_mix16.hpp), written as mixed records: R and
// S = I - T in FP32 register order of the FP32 matrix instruction, sources FP64.
// ---------------------------------------------------------------------------
template <bool REDO>
__device__ __forceinline__ void layer16_task(const FwdArgs& a, long task, const Lane& L,
                                             double* sm) {
  using R = double;
  const int L_ = a.L;
  const int w = (int)(task / L_), l = (int)(task % L_);
  const long gw = (long)(a.w0 + w);
  R* tile = sm + Lds16::tile;
  R* tile2 = sm + Lds16::tile2;
  R* pv = sm + Lds16::pv;
  R* v0 = sm + Lds16::vs;
  R* v1 = v0 + N;
  R* v2 = v1 + N;
  R* v3 = v2 + N;
  R* out = a.lay + task * REC16;

  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  const int mode = a.mode[task];
  const double Btop = a.B ? a.B[gw * (L_ + 1) + lev(a, l)] : 0.0;
  const double Bbot = a.B ? a.B[gw * (L_ + 1) + lev(a, l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (tau > 0.0) ? (Bbot - Btop) / tau : 0.0;
  const int j = L.j;
  const double mu_j = a.mu[j];

  if (tau <= 0.0 || omega_in <= 0.0) {
    // transparent (src/doubling.hpp:52-53) or pure absorption (:56-77)
    double trans = 1.0, su = 0.0, sd = 0.0;
    if (tau > 0.0) {
      const double tex = -tau / mu_j;
      trans = (tex > -200.0) ? exp(tex) : 0.0;
      const double omt = 1.0 - trans;
      const double slope = mu_j * omt - 0.5 * tau * (1.0 + trans);
      su = B_bar * omt + B_d * slope;
      sd = B_bar * omt - B_d * slope;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      out[L.l * 4 + r] = 0.0;
      out[N * N + L.l * 4 + r] = (row<R>(L.h, r) == j) ? trans : 0.0;
    }
    if (L.h == 0) {
      out[2 * N * N + j] = su;
      out[2 * N * N + N + j] = sd;
      if (a.basis) {  // thermal basis: (1 - t, slope) (zero for a transparent layer)
        const double omt = tau > 0.0 ? 1.0 - trans : 0.0;
        const double slope = tau > 0.0 ? mu_j * omt - 0.5 * tau * (1.0 + trans) : 0.0;
        a.basis[task * 2 * N + j] = omt;
        a.basis[task * 2 * N + N + j] = slope;
      }
    }
    if (L.l == 0) a.lflag[task] = 0;
    return;
  }

  Start16<R> st;
  layer16_start<R>(a, task, w, l, gw, L, tau, omega_in, mode, tile, pv, st);
  const int nn = st.nn;
  Mat<R> Rm, Tm;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Tm.v[r] = st.Tv[r];
    Rm.v[r] = st.Rv[r];
  }
  double y = st.y, z = 0.0, sus = st.sus, sds = st.sds, gam = st.gam, gk = st.gk;

  // ---- doubling recursion (src/doubling.hpp:149-209) ----
  for (int k = 0; k < nn; ++k) {
    Mat<R> A, X, Y;
    put<R>(tile, L, Rm);
    if (L.h == 0) {
      v0[j] = sds;
      v1[j] = sus;
    }
    wave_sync();
    mul<R>(tile, Rm, L, Y);  // P = R^2
    double b1 = 0.0, b2 = 0.0;
    if (a.has_solar) {
      double bs[2];
      matvecs<R, 2>(tile, v0, L, bs);  // R s_down_sol, R s_up_sol
      b1 = bs[0];
      b2 = bs[1];
    }
    wave_sync();
    put<R>(tile, L, Y);
    wave_sync();
    if (inverse_ns<R, false>(Y, tile, tile2, A, L)) {  // (I - R^2)^-1
      right_solve_tile<R>(tile, Tm, X, L);              // or TG by a solve
    } else {
      put<R>(tile2, L, Tm);
      wave_sync();
      mul<R>(tile2, A, L, X);  // TG = T (I - R^2)^-1
    }
    wave_sync();
    put<R>(tile2, L, X);
    const double zpgy = z + gk * y;
    if (L.h == 0) {
      v0[j] = zpgy;
      v1[j] = y;
      v2[j] = b1 + gam * sus;
      v3[j] = gam * b2 + sds;
    }
    wave_sync();
    mul<R>(tile2, Rm, L, Y);  // TGR = TG R
    mul<R>(tile2, Tm, L, A);  // T_new = TG T
    double a1, a3, c1 = 0.0, c2 = 0.0;
    if (a.has_solar) {
      double s4[4];
      matvecs<R, 4>(tile2, v0, L, s4);
      a1 = s4[0];
      a3 = s4[1];
      c1 = s4[2];
      c2 = s4[3];
    } else {
      double s2[2];
      matvecs<R, 2>(tile2, v0, L, s2);
      a1 = s2[0];
      a3 = s2[1];
    }
    put<R>(tile, L, Y);
    wave_sync();
    mul_add<R>(tile, Tm, Rm, L, Rm);  // R_new = R + TGR T
    double s2[2];
    matvecs<R, 2>(tile, v0, L, s2);
    const double a2 = s2[0], a4 = s2[1];
    wave_sync();
    Tm = A;
    const double znew = (a1 - a2) + z - gk * y;
    const double ynew = a3 + a4 + y;
    z = znew;
    y = ynew;
    if (a.has_solar) {
      sus = c1 + sus;
      sds = c2 + gam * sds;
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // ---- layer assembly (src/doubling.hpp:211-229), sources combined ----
  // REDO: the FP64 record stays FP64 (flag 2) in a mixed chunk: an unstable
  // start leaves entries far above one in R and T whose FP32 rounding the
  // adding would amplify
  gstore<R>(out, L, Rm);
  gstore<R>(out + N * N, L, Tm);
  if (L.h == 0) {
    out[2 * N * N + j] = (y * B_bar + z * B_d) + sus;
    out[2 * N * N + N + j] = (y * B_bar - z * B_d) + sds;
    if (a.basis) {
      a.basis[task * 2 * N + j] = y;
      a.basis[task * 2 * N + N + j] = z;
    }
  }
  if (L.l == 0) a.lflag[task] = REDO ? 2 : 1;
}

// UNPAIRED: only the layers pair mode leaves (orders up to 8 on reverse-pass
// workspaces: a.pflag, adrt_pair16.hpp).  4 waves per SIMD (122 registers, no
// spills): 120.9 ms on the 65536 x 100 workload against 121.4 - 122.5 ms at 5
// (96 registers, 22 spilled).
template <bool REDO, bool UNPAIRED = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) layer16_kernel(FwdArgs a) {
  __shared__ double smem[4 * Lds16::size];
  const int wv = wave_index();
  const long task = (long)blockIdx.x * 4 + wv;
  if (task >= (long)a.W * a.L) return;
  if constexpr (REDO) {
    if (!a.redo[task]) return;
  }
  if constexpr (UNPAIRED) {
    if (a.pflag[task]) return;
  }
  const Lane L;
  layer16_task<REDO>(a, task, L, smem + wv * Lds16::size);
}

// ---------------------------------------------------------------------------
// adding and interface intensities, one wave per wavenumber
//
// The interface intensities solve the two-stream system
//   I_up(l)   = R_l I_dn(l) + T_l I_up(l+1) + s_up(l)
//   I_dn(l+1) = T_l I_dn(l) + R_l I_up(l+1) + s_dn(l)
// with I_dn(0) given and, at the bottom, I_up(L) = R_s I_dn(L) + s_s (the
// surface) or I_up(L) given.  The reference solves it by adding: composites
// of all layers above and of all layers below every interface (two sweeps,
// two inverses per composite) and a solve of the two composites at each
// interface (src/solver.cpp:24-73, 560-700; src/adding.hpp:31-258).  Here the
// same system is solved by one sweep of the base composites only -- Rb(l),
// the reflectance of everything below interface l seen from above, and
// sb(l), the upward intensity leaving it with nothing incident from above
// (its sources and the bottom boundary's intensity carried through it):
//   X_l = (I - R_l Rb(l+1))^-1,   G_l = X_l T_l,
//   d_l = X_l (R_l sb(l+1) + s_dn(l)),
//   Rb(l) = R_l + T_l Rb(l+1) G_l,   sb(l) = s_up(l) + T_l (Rb(l+1) d_l + sb(l+1)),
// from Rb(L) = R_s, sb(L) = s_s (or Rb(L) = 0, sb(L) = the given I_up(L));
// then one downward sweep of matrix-vector products
//   I_dn(l+1) = G_l I_dn(l) + d_l,   I_up(l) = Rb(l) I_dn(l) + sb(l),
// whose I_dn(l+1), I_up(l+1) satisfy both equations of layer l for the
// I_dn(l) above it and the base below it.  One inverse and four matrix
// products per scattering layer, against six inverses and about twenty
// products for the two composite sweeps and the interface solves.  The
// transposed system has the same form with R_l^T, T_l^T, the output
// cotangents as sources and the roles of up and down exchanged, so the
// reverse pass runs this kernel with ADJ = true.
// ---------------------------------------------------------------------------
struct Ws16 {
  double *tile, *tile2, *pv, *v0, *v1, *v2, *v3;
};

// cotangents of I_up(l) and I_dn(l) (element j) from the output cotangents:
// flux_up = scaling 2 pi sum w mu I_up, flux_down = 2 pi sum w mu I_dn,
// J = 1/2 sum w (I_up + I_dn) + beam, divergence = 4 pi (1 - omega)(J - B)
__device__ __forceinline__ void ibar16(const FwdArgs& a, long gw, int l, double mu_j, double wt_j,
                                       double& up, double& dn) {
  const long o = gw * (a.L + 1) + lev(a, l);
  const double cfu = a.cfu ? a.cfu[o] : 0.0;
  const double cfd = a.cfd ? a.cfd[o] : 0.0;
  double cJ = a.cJ ? a.cJ[o] : 0.0;
  if (a.cdiv) {
    const int lyr = (l == 0) ? 0 : l - 1;
    cJ += a.cdiv[o] * (4.0 * kPi * (1.0 - a.ssa[gw * a.L + lin(a, lyr)]));
  }
  const double c = 2.0 * kPi * wt_j * mu_j;
  up = cfu * a.scaling * c + 0.5 * wt_j * cJ;
  dn = cfd * c + 0.5 * wt_j * cJ;
}

// layer l of the forward system, or of its transpose (ADJ): transposed R and
// T, source up = cotangent of I_dn(l), source down = cotangent of I_up(l+1).
// The matrices are read on demand so that they are live only where used.
// F32: records of the mixed-precision layer kernel (R and S = I - T in FP32
// register order of the FP32 matrix instruction, rows 4h + r; sources FP64
// after them); T is rebuilt as I - S in FP64.  Layers the FP64 kernel redid
// (flag 2) keep the FP64 record.
template <bool ADJ, bool F32>
struct LayerRef {
  const FwdArgs& a;
  const double* rec;
  long gw;
  double mu_j, wt_j;
  int scat;
  __device__ bool narrow() const { return F32 && scat != 2; }
  // element (row<double>(h, r), j) of a stored N x N matrix, or of its
  // transpose; eye_minus: the record holds I - M
  __device__ static void load(const double* m, const Lane& L, bool trans, bool eye_minus,
                              Mat<double>& M) {
    if constexpr (F32) {
      const float* f = (const float*)m;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row<double>(L.h, r);
        // FP32 element (p, c) sits in lane c + 16 (p >> 2), register p & 3
        const double v = trans ? (double)f[(i + 16 * (L.j >> 2)) * 4 + (L.j & 3)]
                               : (double)f[(L.j + 16 * (i >> 2)) * 4 + (i & 3)];
        M.v[r] = eye_minus ? ((i == L.j) ? 1.0 : 0.0) - v : v;
      }
    } else {
      if (trans) gload_t<double>(m, L, M);
      else gload<double>(m, L, M);
    }
  }
  __device__ static void load64(const double* m, const Lane& L, bool trans, Mat<double>& M) {
    if (trans) gload_t<double>(m, L, M);
    else gload<double>(m, L, M);
  }
  __device__ const double* tmat() const {
    return F32 ? (const double*)((const float*)rec + N * N) : rec + N * N;
  }
  __device__ void mats(const Lane& L, Mat<double>& R, Mat<double>& T) const {
    if (narrow()) {
      load(rec, L, ADJ, false, R);
      load(tmat(), L, ADJ, true, T);
    } else {
      load64(rec, L, ADJ, R);
      load64(rec + N * N, L, ADJ, T);
    }
  }
  __device__ void refl(const Lane& L, Mat<double>& R) const {
    if (narrow()) load(rec, L, ADJ, false, R);
    else load64(rec, L, ADJ, R);
  }
  __device__ void trans(const Lane& L, Mat<double>& T) const {
    if (narrow()) load(tmat(), L, ADJ, true, T);
    else load64(rec + N * N, L, ADJ, T);
  }
  __device__ void sources(const Lane& L, int l, double& su, double& sd) const {
    if constexpr (ADJ) {
      double u, d;
      ibar16(a, gw, l, mu_j, wt_j, u, d);
      su = d;
      ibar16(a, gw, l + 1, mu_j, wt_j, u, d);
      sd = u;
    } else {
      const double* src = rec + (narrow() ? N * N : 2 * N * N);
      su = src[L.j];
      sd = src[N + L.j];
    }
  }
};

template <bool ADJ, bool F32>
__device__ __forceinline__ LayerRef<ADJ, F32> layer_ref(const FwdArgs& a, const double* lay,
                                                        const int* lflag, long gw, int l,
                                                        double mu_j, double wt_j) {
  return LayerRef<ADJ, F32>{a, lay + (long)l * REC16, gw, mu_j, wt_j, lflag[l]};
}

// the base below an interface: Rb (lane order; all zero unless scat) and sb
// (element j)
struct Base16 {
  Mat<double> R;
  double s;
  int scat;
};

// Layer l (y) on top of the base B: stores at o what the downward sweep needs
// at layer l -- G_l (the diagonal of T_l for a non-scattering layer), d_l and
// the base below, Rb(l+1), sb(l+1) -- and replaces B by the base below
// interface l.
template <class LR>
__device__ void base16_add(Base16& B, const LR& y, int l, double* o, const Lane& L,
                           const Ws16& s) {
  using R = double;
  const int j = L.j;
  R su, sd;
  y.sources(L, l, su, sd);
  if (B.scat) gstore<R>(o + RB_R, L, B.R);
  if (L.h == 0) {
    o[RB_S + j] = B.s;
    if (j == 0) o[RB_F] = B.scat ? 1.0 : 0.0;
  }
  if (!y.scat) {
    // T_l = diag(t), R_l = 0: X = I, G = T_l, d = s_dn,
    // Rb(l) = T_l Rb T_l, sb(l) = s_up + T_l (Rb s_dn + sb)
    Mat<R> T;
    y.trans(L, T);
    const R tj = diag<R>(T, L);
    R rsd = 0.0;
    if (B.scat) {
      put<R>(s.tile, L, B.R);
      if (L.h == 0) {
        s.v0[j] = tj;
        s.v1[j] = sd;
      }
      wave_sync();
      rsd = matvec<R>(s.tile, s.v1, L);
      R tv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) tv[r] = s.v0[row<R>(L.h, r)];
      wave_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) B.R.v[r] = tv[r] * B.R.v[r] * tj;
    }
    if (L.h == 0) {
      o[RB_G + j] = tj;
      o[RB_D + j] = sd;
    }
    B.s = su + tj * (rsd + B.s);
    return;
  }
  if (!B.scat) {
    // nothing below reflects: X = I, G = T_l, d = R_l sb + s_dn, Rb(l) = R_l,
    // sb(l) = s_up + T_l sb
    Mat<R> T;
    y.mats(L, B.R, T);
    put<R>(s.tile, L, B.R);
    put<R>(s.tile2, L, T);
    if (L.h == 0) s.v0[j] = B.s;
    wave_sync();
    const R d = matvec<R>(s.tile, s.v0, L) + sd;
    const R ts = matvec<R>(s.tile2, s.v0, L);
    wave_sync();
    gstore<R>(o + RB_G, L, T);
    if (L.h == 0) o[RB_D + j] = d;
    B.s = su + ts;
    B.scat = 1;
    return;
  }
  Mat<R> Rl, P, X, G, T;
  y.refl(L, Rl);
  put<R>(s.tile, L, Rl);
  if (L.h == 0) s.v0[j] = B.s;
  wave_sync();
  mul<R>(s.tile, B.R, L, P);                      // P = R_l Rb(l+1)
  const R v = matvec<R>(s.tile, s.v0, L) + sd;    // R_l sb(l+1) + s_dn
  wave_sync();
  put<R>(s.tile, L, P);
  if (L.h == 0) s.v1[j] = v;
  wave_sync();
  R d;
  if (inverse_ns<R, false>(P, s.tile, s.tile2, X, L)) {
    // I - P close to singular, or the iteration slow: G and d by LU solves
    y.trans(L, T);
    left_solve_mat_tile<R>(s.tile, T, G, s.tile2, L);
    d = left_solve_tile<R>(s.tile, v, L);
  } else {
    y.trans(L, T);
    put<R>(s.tile2, L, X);
    wave_sync();
    mul<R>(s.tile2, T, L, G);                     // G = X T_l
    d = matvec<R>(s.tile2, s.v1, L);              // d = X (R_l sb + s_dn)
  }
  wave_sync();
  gstore<R>(o + RB_G, L, G);
  if (L.h == 0) o[RB_D + j] = d;
  put<R>(s.tile, L, B.R);
  if (L.h == 0) s.v2[j] = d;
  wave_sync();
  Mat<R> M;
  mul<R>(s.tile, G, L, M);                        // Rb(l+1) G
  const R rd = matvec<R>(s.tile, s.v2, L);        // Rb(l+1) d
  wave_sync();
  put<R>(s.tile, L, T);
  if (L.h == 0) s.v3[j] = rd + B.s;
  wave_sync();
  y.refl(L, Rl);
  mul_add<R>(s.tile, M, Rl, L, B.R);              // Rb(l) = R_l + T_l Rb(l+1) G
  const R ts = matvec<R>(s.tile, s.v3, L);
  wave_sync();
  B.s = su + ts;
}

template <bool ADJ, bool F32 = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
adding16_kernel(FwdArgs a) {
  using R = double;
  __shared__ R smem[4 * Lds16::size];
  const int wv = wave_index();
  const int w = blockIdx.x * 4 + wv;
  if (w >= a.W || (a.ucol && a.ucol[w])) return;
  const Lane L;
  const int j = L.j;
  const long gw = (long)(a.w0 + w);
  const int Lay = a.L;
  R* sm = smem + wv * Lds16::size;
  Ws16 s{sm + Lds16::tile, sm + Lds16::tile2, sm + Lds16::pv,
         sm + Lds16::vs, sm + Lds16::vs + N, sm + Lds16::vs + 2 * N, sm + Lds16::vs + 3 * N};
  const double* lay = a.lay + (long)w * Lay * REC16;
  const int* lflag = a.lflag + (long)w * Lay;
  double* rbw = a.rb + (long)w * Lay * RBR16;
  const double mu_j = a.mu[j], wt_j = a.wt[j];
  const double* Bw = a.B ? a.B + gw * (Lay + 1) : nullptr;
  auto Blev = [&](int l) -> double { return Bw ? Bw[lev(a, l)] : 0.0; };
  double* vec = ADJ ? (a.Av ? a.Av + (long)w * (Lay + 1) * 2 * N : nullptr)
                    : (a.Iv ? a.Iv + (long)w * (Lay + 1) * 2 * N : nullptr);

  // boundary conditions (src/solver.cpp:385-419, 516-551, 585-603)
  const double Bs = a.surf_w ? a.surf_w[gw] : a.surf_em;
  const double Btop = a.top_w ? a.top_w[gw] : a.top_em;
  const double A = a.albedo;
  const int has_surface = !a.diffusion && (A > 0.0 || Bs > 0.0);
  const double tau_tot = a.tau_cum[(long)w * (Lay + 1) + Lay];
  double Itd, Ibu = 0.0;
  if constexpr (ADJ) {
    double u, d;
    ibar16(a, gw, 0, mu_j, wt_j, u, d);
    Itd = u;  // transposed system: the top boundary carries the cotangent of I_up(0)
    ibar16(a, gw, Lay, mu_j, wt_j, u, d);
    if (!has_surface) Ibu = d;  // ... and a fixed I_up(L) the cotangent of I_dn(L)
  } else {
    Itd = Btop;
    if (a.diffusion) {
      const double Bb = Blev(Lay);
      const double dt = a.tau_s[(long)w * Lay + Lay - 1];
      const double dBdt = (dt > 0.0) ? (Bb - Blev(Lay - 1)) / dt : 0.0;
      Ibu = Bb + mu_j * dBdt;
    } else if (!has_surface) {
      Ibu = Bs;
    }
  }

  // ---- the base below the last layer (src/solver.cpp:385-419, 516-551, 585-603) ----
  Base16 B;
  if (has_surface) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = row<R>(L.h, r);
      if constexpr (ADJ)  // R_s^T(i, j) = R_s(j, i)
        B.R.v[r] = (j < a.n_real) ? 2.0 * A * a.mu[i] * a.wt[i] * a.xfac : 0.0;
      else
        B.R.v[r] = (i < a.n_real) ? 2.0 * A * mu_j * wt_j * a.xfac : 0.0;
    }
    if constexpr (ADJ) {
      double u, d;
      ibar16(a, gw, Lay, mu_j, wt_j, u, d);
      B.s = d;
    } else {
      double ssol = 0.0;
      if (a.has_solar && A > 0.0) ssol = (A / kPi) * a.F0 * a.mu0 * exp(-tau_tot / a.mu0);
      B.s = (1.0 - A) * Bs + ssol;
    }
    B.scat = A > 0.0;
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) B.R.v[r] = 0.0;
    B.s = Ibu;
    B.scat = 0;
  }

  // ---- base sweep, bottom to top ----
  for (int l = Lay - 1; l >= 0; --l)
    base16_add(B, layer_ref<ADJ, F32>(a, lay, lflag, gw, l, mu_j, wt_j), l,
               rbw + (long)l * RBR16, L, s);

  auto emit = [&](int l, double Iup, double Idn) {
    if (vec && L.h == 0) {
      vec[(long)l * 2 * N + j] = Iup;
      vec[(long)l * 2 * N + N + j] = Idn;
    }
    if constexpr (!ADJ) {
      const double c = 2.0 * kPi * wt_j * mu_j;
      const double fu = colsum<R>(c * Iup, L, s.pv);
      const double fdn = colsum<R>(c * Idn, L, s.pv);
      double J = colsum<R>(0.5 * wt_j * (Iup + Idn), L, s.pv);
      if (L.l == 0) {
        const long o = gw * (Lay + 1) + lev(a, l);
        double fdir = 0.0;
        if (a.has_solar) {
          fdir = (l == 0) ? a.F0 * a.mu0
                          : a.F0 * a.mu0 * exp(-a.tau_cum[(long)w * (Lay + 1) + l] / a.mu0);
          J += fdir * (1.0 / (4.0 * kPi * a.mu0));
        }
        a.fu[o] = fu * a.scaling;
        a.fd[o] = fdn;
        a.mi[o] = J;
        a.fdir[o] = fdir;
        if (a.fdiv) {
          const int lyr = (l == 0) ? 0 : l - 1;
          const double om = a.ssa[gw * Lay + lin(a, lyr)];
          a.fdiv[o] = 4.0 * kPi * (1.0 - om) * (J - Blev(l));
        }
      }
    }
  };

  // ---- downward sweep (the interface intensities of src/solver.cpp:628-678) ----
  double Idn = Itd, Iup = B.s;
  if (B.scat) {
    put<R>(s.tile, L, B.R);
    if (L.h == 0) s.v0[j] = Itd;
    wave_sync();
    Iup += matvec<R>(s.tile, s.v0, L);
    wave_sync();
  }
  emit(0, Iup, Idn);
  for (int l = 0; l < Lay; ++l) {
    const double* o = rbw + (long)l * RBR16;
    const double d = o[RB_D + j], sbn = o[RB_S + j];
    const bool bscat = o[RB_F] != 0.0;
    if (lflag[l]) {
      Mat<R> G;
      gload<R>(o + RB_G, L, G);
      put<R>(s.tile, L, G);
      if (L.h == 0) s.v0[j] = Idn;
      wave_sync();
      Idn = matvec<R>(s.tile, s.v0, L) + d;
      wave_sync();
    } else {
      Idn = o[RB_G + j] * Idn + d;
    }
    Iup = sbn;
    if (bscat) {
      Mat<R> Rb;
      gload<R>(o + RB_R, L, Rb);
      put<R>(s.tile, L, Rb);
      if (L.h == 0) s.v0[j] = Idn;
      wave_sync();
      Iup += matvec<R>(s.tile, s.v0, L);
      wave_sync();
    }
    emit(l + 1, Iup, Idn);
  }
}

// ---------------------------------------------------------------------------
// adding in the reference's form, for the columns the base sweep does not
// hold to the reference: columns in which two scattering layers with an
// unstable doubling start follow each other (non-scattering layers between
// them aside).  Their records are far from physical (entries of R and T up to
// 1e8 where the reference's start amplifies its rounding), and at the
// interface between two such layers I_up = Rb I_dn + sb cancels terms of
// that size; the reference's composites of the layers above and below each
// interface and its solve there (src/adding.hpp:69-258, src/solver.cpp:24-73,
// 560-700) land closer to the reference's own result.  The same system,
// solved as the reference does: composites C (R_ab, R_ba, T_ab, T_ba, sources)
// of everything below each interface (bottom-up), then of everything above it
// (top-down), and at each interface the two 16 x 16 solves for I_up and I_dn.
// ---------------------------------------------------------------------------
struct Comp16 {
  Mat<double> Rab, Rba, Tab, Tba;
  double su, sd;
  int scat;
};


// rbase step: new = add(top = layer y, bot = C)  (src/adding.hpp:246-258)
template <class LR>
__device__ void add16_on_top(Comp16& C, const LR& y, int l, const Lane& L,
                             const Ws16& s) {
  using R = double;
  const int j = L.j;
  R tsu, tsd;
  y.sources(L, l, tsu, tsd);
  Mat<R> Rt, Tt, A, X;
  if (!y.scat) {
    // NonscatTop (src/adding.hpp:117-174): t = diag(top.T_ab)
    y.trans(L, Tt);
    const R tj = diag<R>(Tt, L);
    if (L.h == 0) {
      s.v0[j] = tj;
      s.v1[j] = tsd;
    }
    put<R>(s.tile, L, C.Rab);
    put<R>(s.tile2, L, C.Tab);
    wave_sync();
    const R rbc = matvec<R>(s.tile, s.v1, L);   // bot.R_ab s1_down
    const R tbc = matvec<R>(s.tile2, s.v1, L);  // bot.T_ab s1_down
    R tv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) tv[r] = s.v0[row<R>(L.h, r)];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      C.Rab.v[r] = tv[r] * C.Rab.v[r] * tj;
      C.Tab.v[r] = C.Tab.v[r] * tj;
      C.Tba.v[r] = tv[r] * C.Tba.v[r];
    }
    const R nsu = tsu + tj * (C.su + rbc);
    const R nsd = C.sd + tbc;
    C.su = nsu;
    C.sd = nsd;
    return;
  }
  y.mats(L, Rt, Tt);
  if (!C.scat) {
    // NonscatBot (src/adding.hpp:181-239): t = diag(bot.T_ab)
    const R tj = diag<R>(C.Tab, L);
    if (L.h == 0) {
      s.v0[j] = tj;
      s.v1[j] = C.su;
    }
    put<R>(s.tile, L, Tt);
    put<R>(s.tile2, L, Rt);
    wave_sync();
    const R tba = matvec<R>(s.tile, s.v1, L);   // top.T_ba s2_up
    const R rba = matvec<R>(s.tile2, s.v1, L);  // top.R_ba s2_up
    R tv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) tv[r] = s.v0[row<R>(L.h, r)];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      C.Rab.v[r] = Rt.v[r];
      C.Rba.v[r] = tv[r] * Rt.v[r] * tj;
      C.Tab.v[r] = tv[r] * Tt.v[r];
      C.Tba.v[r] = Tt.v[r] * tj;
    }
    const R nsu = tsu + tba;
    const R nsd = C.sd + tj * (tsd + rba);
    C.su = nsu;
    C.sd = nsd;
    C.scat = 1;
    return;
  }
  // General (src/adding.hpp:69-110), the two inverses by inverse_ns
  put<R>(s.tile, L, C.Rab);
  put<R>(s.tile2, L, Rt);
  if (L.h == 0) {
    s.v0[j] = tsd;
    s.v1[j] = C.su;
  }
  wave_sync();
  mul<R>(s.tile, Rt, L, A);                               // P1 = bot.R_ab top.R_ba
  const R rhs_up = C.su + matvec<R>(s.tile, s.v0, L);     // s2_up + R_bc s1_down
  const R rhs_dn = tsd + matvec<R>(s.tile2, s.v1, L);     // s1_down + R_ba s2_up
  Mat<R> B;
  mul<R>(s.tile2, C.Rab, L, B);                           // P2 = top.R_ba bot.R_ab
  wave_sync();
  put<R>(s.tile, L, A);
  wave_sync();
  if (inverse_ns<R, false>(A, s.tile, s.tile2, X, L)) {   // A1^-1 = (I - P1)^-1
    right_solve_tile<R>(s.tile, Tt, X, L);                // or D1 by a solve
    put<R>(s.tile, L, B);
  } else {
    put<R>(s.tile, L, B);
    put<R>(s.tile2, L, Tt);
    wave_sync();
    mul<R>(s.tile2, X, L, X);                             // D1 = top.T_ba A1^-1
  }
  wave_sync();
  if (inverse_ns<R, false>(B, s.tile, s.tile2, A, L)) {   // A2^-1 = (I - P2)^-1
    right_solve_tile<R>(s.tile, C.Tab, A, L);             // or D2 by a solve
    C.Tab = A;
  } else {
    put<R>(s.tile, L, C.Tab);
    wave_sync();
    mul<R>(s.tile, A, L, C.Tab);                          // D2 = bot.T_ab A2^-1
  }
  wave_sync();
  if (L.h == 0) {
    s.v2[j] = rhs_up;
    s.v3[j] = rhs_dn;
  }
  put<R>(s.tile, L, C.Tab);  // D2
  put<R>(s.tile2, L, X);     // D1
  wave_sync();
  mul<R>(s.tile, Rt, L, A);                                // D2 top.R_ba
  const R nsd = C.sd + matvec<R>(s.tile, s.v3, L);
  mul<R>(s.tile, Tt, L, C.Tab);                           // T_ab = D2 top.T_ab
  mul<R>(s.tile2, C.Rab, L, B);                           // D1 bot.R_ab
  const R nsu = tsu + matvec<R>(s.tile2, s.v2, L);
  mul<R>(s.tile2, C.Tba, L, X);                           // D1 bot.T_ba (new T_ba)
  wave_sync();
  put<R>(s.tile, L, A);
  put<R>(s.tile2, L, B);
  wave_sync();
  mul_add<R>(s.tile, C.Tba, C.Rba, L, C.Rba);             // R_ba = bot.R_ba + (D2 R_ba) T_ba
  mul_add<R>(s.tile2, Tt, Rt, L, C.Rab);                  // R_ab = top.R_ab + (D1 R_ab) T_ab
  wave_sync();
  C.Tba = X;
  C.su = nsu;
  C.sd = nsd;
  C.scat = 1;
}

// rtop step: new = add(top = C, bot = layer y)
template <class LR>
__device__ void add16_below(Comp16& C, const LR& y, int l, const Lane& L,
                            const Ws16& s) {
  using R = double;
  const int j = L.j;
  R bsu, bsd;
  y.sources(L, l, bsu, bsd);
  Mat<R> Rb, Tb, A, X;
  y.mats(L, Rb, Tb);
  if (!C.scat) {
    // NonscatTop with top = C: t = diag(C.T_ab)
    const R tj = diag<R>(C.Tab, L);
    if (L.h == 0) {
      s.v0[j] = tj;
      s.v1[j] = C.sd;
    }
    put<R>(s.tile, L, Rb);
    put<R>(s.tile2, L, Tb);
    wave_sync();
    const R rbc = matvec<R>(s.tile, s.v1, L);   // bot.R_ab s1_down
    const R tbc = matvec<R>(s.tile2, s.v1, L);  // bot.T_ab s1_down
    R tv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) tv[r] = s.v0[row<R>(L.h, r)];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      C.Rab.v[r] = tv[r] * Rb.v[r] * tj;
      C.Rba.v[r] = Rb.v[r];
      C.Tab.v[r] = Tb.v[r] * tj;
      C.Tba.v[r] = tv[r] * Tb.v[r];
    }
    const R nsu = C.su + tj * (bsu + rbc);
    const R nsd = bsd + tbc;
    C.su = nsu;
    C.sd = nsd;
    C.scat = y.scat;
    return;
  }
  if (!y.scat) {
    // NonscatBot with bot = layer: t = diag(layer T)
    const R tj = diag<R>(Tb, L);
    if (L.h == 0) {
      s.v0[j] = tj;
      s.v1[j] = bsu;
    }
    put<R>(s.tile, L, C.Tba);
    put<R>(s.tile2, L, C.Rba);
    wave_sync();
    const R tba = matvec<R>(s.tile, s.v1, L);   // top.T_ba s2_up
    const R rba = matvec<R>(s.tile2, s.v1, L);  // top.R_ba s2_up
    R tv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) tv[r] = s.v0[row<R>(L.h, r)];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      C.Rba.v[r] = tv[r] * C.Rba.v[r] * tj;
      C.Tab.v[r] = tv[r] * C.Tab.v[r];
      C.Tba.v[r] = C.Tba.v[r] * tj;
    }
    const R nsu = C.su + tba;
    const R nsd = bsd + tj * (C.sd + rba);
    C.su = nsu;
    C.sd = nsd;
    return;
  }
  // General with top = C, bot = layer; the inverses as in add16_on_top
  put<R>(s.tile, L, Rb);
  put<R>(s.tile2, L, C.Rba);
  if (L.h == 0) {
    s.v0[j] = C.sd;
    s.v1[j] = bsu;
  }
  wave_sync();
  mul<R>(s.tile, C.Rba, L, A);                            // P1 = bot.R_ab top.R_ba
  const R rhs_up = bsu + matvec<R>(s.tile, s.v0, L);      // s2_up + R_bc s1_down
  const R rhs_dn = C.sd + matvec<R>(s.tile2, s.v1, L);    // s1_down + R_ba s2_up
  Mat<R> B;
  mul<R>(s.tile2, Rb, L, B);                              // P2 = top.R_ba bot.R_ab
  wave_sync();
  put<R>(s.tile, L, A);
  wave_sync();
  if (inverse_ns<R, false>(A, s.tile, s.tile2, X, L)) {   // A1^-1
    right_solve_tile<R>(s.tile, C.Tba, X, L);             // or D1 by a solve
    C.Tba = X;
    put<R>(s.tile, L, B);
  } else {
    put<R>(s.tile, L, B);
    put<R>(s.tile2, L, C.Tba);
    wave_sync();
    mul<R>(s.tile2, X, L, C.Tba);                         // D1 = top.T_ba A1^-1
  }
  wave_sync();
  if (inverse_ns<R, false>(B, s.tile, s.tile2, A, L)) {   // A2^-1
    right_solve_tile<R>(s.tile, Tb, X, L);                // or D2 by a solve
  } else {
    put<R>(s.tile, L, Tb);
    wave_sync();
    mul<R>(s.tile, A, L, X);                              // D2 = bot.T_ab A2^-1
  }
  wave_sync();
  if (L.h == 0) {
    s.v2[j] = rhs_up;
    s.v3[j] = rhs_dn;
  }
  put<R>(s.tile, L, C.Tba);  // D1
  put<R>(s.tile2, L, X);     // D2
  wave_sync();
  mul<R>(s.tile, Rb, L, A);                                // D1 bot.R_ab
  const R nsu = C.su + matvec<R>(s.tile, s.v2, L);
  mul<R>(s.tile, Tb, L, C.Tba);                           // T_ba = D1 bot.T_ba
  mul<R>(s.tile2, C.Rba, L, B);                           // D2 top.R_ba
  const R nsd = bsd + matvec<R>(s.tile2, s.v3, L);
  mul<R>(s.tile2, C.Tab, L, X);                           // T_ab = D2 top.T_ab
  wave_sync();
  put<R>(s.tile, L, A);
  put<R>(s.tile2, L, B);
  wave_sync();
  mul_add<R>(s.tile, C.Tab, C.Rab, L, C.Rab);             // R_ab = top.R_ab + (D1 R_ab) T_ab
  mul_add<R>(s.tile2, Tb, Rb, L, C.Rba);                  // R_ba = bot.R_ba + (D2 R_ba) T_ba
  wave_sync();
  C.Tab = X;
  C.su = nsu;
  C.sd = nsd;
  C.scat = 1;
}

template <bool ADJ, bool F32 = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
adding16_composite_kernel(FwdArgs a) {
  using R = double;
  __shared__ R smem[4 * Lds16::size];
  const int wv = wave_index();
  int w = blockIdx.x * 4 + wv;
  if (!composite_column(a, w)) return;
  const Lane L;
  const int j = L.j;
  const long gw = (long)(a.w0 + w);
  const int Lay = a.L;
  R* sm = smem + wv * Lds16::size;
  Ws16 s{sm + Lds16::tile, sm + Lds16::tile2, sm + Lds16::pv,
         sm + Lds16::vs, sm + Lds16::vs + N, sm + Lds16::vs + 2 * N, sm + Lds16::vs + 3 * N};
  const double* lay = a.lay + (long)w * Lay * REC16;
  const int* lflag = a.lflag + (long)w * Lay;
  double* rbw = a.rb + (long)w * Lay * RBR16;
  const double mu_j = a.mu[j], wt_j = a.wt[j];
  const double* Bw = a.B ? a.B + gw * (Lay + 1) : nullptr;
  auto Blev = [&](int l) -> double { return Bw ? Bw[lev(a, l)] : 0.0; };
  double* vec = ADJ ? (a.Av ? a.Av + (long)w * (Lay + 1) * 2 * N : nullptr)
                    : (a.Iv ? a.Iv + (long)w * (Lay + 1) * 2 * N : nullptr);

  // boundary conditions (src/solver.cpp:385-419, 516-551, 585-603)
  const double Bs = a.surf_w ? a.surf_w[gw] : a.surf_em;
  const double Btop = a.top_w ? a.top_w[gw] : a.top_em;
  const double A = a.albedo;
  const int has_surface = !a.diffusion && (A > 0.0 || Bs > 0.0);
  const int ltot = Lay + (has_surface ? 1 : 0);
  const double tau_tot = a.tau_cum[(long)w * (Lay + 1) + Lay];
  double Itd, Ibu = 0.0;
  if constexpr (ADJ) {
    double u, d;
    ibar16(a, gw, 0, mu_j, wt_j, u, d);
    Itd = u;  // transposed system: the top boundary carries the cotangent of I_up(0)
    ibar16(a, gw, Lay, mu_j, wt_j, u, d);
    if (!has_surface) Ibu = d;  // ... and a fixed I_up(L) the cotangent of I_dn(L)
  } else {
    Itd = Btop;
    if (a.diffusion) {
      const double Bb = Blev(Lay);
      const double dt = a.tau_s[(long)w * Lay + Lay - 1];
      const double dBdt = (dt > 0.0) ? (Bb - Blev(Lay - 1)) / dt : 0.0;
      Ibu = Bb + mu_j * dBdt;
    } else if (!has_surface) {
      Ibu = Bs;
    }
  }

  Comp16 C;
  // ---- bottom-up sweep (src/solver.cpp:560-572) ----
  if (has_surface) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = row<R>(L.h, r);
      if constexpr (ADJ)  // R_s^T(i, j) = R_s(j, i)
        C.Rab.v[r] = (j < a.n_real) ? 2.0 * A * a.mu[i] * a.wt[i] * a.xfac : 0.0;
      else
        C.Rab.v[r] = (i < a.n_real) ? 2.0 * A * mu_j * wt_j * a.xfac : 0.0;
      C.Rba.v[r] = C.Rab.v[r];
      C.Tab.v[r] = 0.0;
      C.Tba.v[r] = 0.0;
    }
    if constexpr (ADJ) {
      double u, d;
      ibar16(a, gw, Lay, mu_j, wt_j, u, d);
      C.su = d;
    } else {
      double ssol = 0.0;
      if (a.has_solar && A > 0.0) ssol = (A / kPi) * a.F0 * a.mu0 * exp(-tau_tot / a.mu0);
      C.su = (1.0 - A) * Bs + ssol;
    }
    C.sd = 0.0;
    C.scat = A > 0.0;
  } else {
    const auto y = layer_ref<ADJ, F32>(a, lay, lflag, gw, Lay - 1, mu_j, wt_j);
    y.mats(L, C.Rab, C.Tab);
    C.Rba = C.Rab;
    C.Tba = C.Tab;
    y.sources(L, Lay - 1, C.su, C.sd);
    C.scat = y.scat;
  }
  for (int n = 1;; ++n) {  // C == rbase[n]
    if (n < ltot) {
      double* o = rbw + (long)(n - 1) * RBR16;
      gstore<R>(o, L, C.Rab);
      gstore<R>(o + N * N, L, C.Tba);
      if (L.h == 0) o[2 * N * N + j] = C.su;
    }
    if (n == ltot) break;
    const int k = ltot - 1 - n;
    add16_on_top(C, layer_ref<ADJ, F32>(a, lay, lflag, gw, k, mu_j, wt_j), k, L, s);
  }

  auto emit = [&](int l, double Iup, double Idn) {
    if (vec && L.h == 0) {
      vec[(long)l * 2 * N + j] = Iup;
      vec[(long)l * 2 * N + N + j] = Idn;
    }
    if constexpr (!ADJ) {
      const double c = 2.0 * kPi * wt_j * mu_j;
      const double fu = colsum<R>(c * Iup, L, s.pv);
      const double fdn = colsum<R>(c * Idn, L, s.pv);
      double J = colsum<R>(0.5 * wt_j * (Iup + Idn), L, s.pv);
      if (L.l == 0) {
        const long o = gw * (Lay + 1) + lev(a, l);
        double fdir = 0.0;
        if (a.has_solar) {
          fdir = (l == 0) ? a.F0 * a.mu0
                          : a.F0 * a.mu0 * exp(-a.tau_cum[(long)w * (Lay + 1) + l] / a.mu0);
          J += fdir * (1.0 / (4.0 * kPi * a.mu0));
        }
        a.fu[o] = fu * a.scaling;
        a.fd[o] = fdn;
        a.mi[o] = J;
        a.fdir[o] = fdir;
        if (a.fdiv) {
          const int lyr = (l == 0) ? 0 : l - 1;
          const double om = a.ssa[gw * Lay + lin(a, lyr)];
          a.fdiv[o] = 4.0 * kPi * (1.0 - om) * (J - Blev(l));
        }
      }
    }
  };

  // ---- TOA (src/solver.cpp:628-642) ----
  {
    put<R>(s.tile, L, C.Rab);
    put<R>(s.tile2, L, C.Tba);
    if (L.h == 0) {
      s.v0[j] = Itd;
      s.v1[j] = Ibu;
    }
    wave_sync();
    const double rI = matvec<R>(s.tile, s.v0, L);
    const double tI = matvec<R>(s.tile2, s.v1, L);
    wave_sync();
    emit(0, rI + (tI + C.su), Itd);
  }

  // ---- top-down sweep with interfaces (src/solver.cpp:574-583, 644-678) ----
  {
    const auto y = layer_ref<ADJ, F32>(a, lay, lflag, gw, 0, mu_j, wt_j);
    y.mats(L, C.Rab, C.Tab);
    C.Rba = C.Rab;
    C.Tba = C.Tab;
    y.sources(L, 0, C.su, C.sd);
    C.scat = y.scat;
  }
  for (int l = 1; l <= Lay; ++l) {  // C == rtop[l]
    const int n_base = ltot - l;
    double Iup, Idn;
    if (n_base > 0) {
      const double* o = rbw + (long)(n_base - 1) * RBR16;
      const double sbu = o[2 * N * N + j];
      Mat<R> Rb, Tb, M;
      gload<R>(o, L, Rb);
      gload<R>(o + N * N, L, Tb);
      // top.T_ab I_top_down ; base.T_ba I_bot_up
      put<R>(s.tile, L, C.Tab);
      put<R>(s.tile2, L, Tb);
      if (L.h == 0) {
        s.v0[j] = Itd;
        s.v1[j] = Ibu;
      }
      wave_sync();
      const double tti = matvec<R>(s.tile, s.v0, L);
      const double tbi = matvec<R>(s.tile2, s.v1, L);
      wave_sync();
      // M_up = I - base.R_ab top.R_ba ; terms 2 and 3 (src/solver.cpp:24-46)
      put<R>(s.tile, L, Rb);
      put<R>(s.tile2, L, C.Rba);
      if (L.h == 0) {
        s.v0[j] = tti;
        s.v1[j] = C.sd;
        s.v2[j] = tbi;
        s.v3[j] = sbu;
      }
      wave_sync();
      mul<R>(s.tile, C.Rba, L, M);  // P_up = base.R_ab top.R_ba, M_up = I - P_up
      double t23[2];
      matvecs<R, 2>(s.tile, s.v0, L, t23);
      double u23[2];
      matvecs<R, 2>(s.tile2, s.v2, L, u23);
      Mat<R> M2;
      mul<R>(s.tile2, Rb, L, M2);  // P_dn = top.R_ba base.R_ab, M_dn = I - P_dn
      const double b_up = tbi + t23[0] + t23[1] + sbu;
      const double b_dn = tti + u23[0] + u23[1] + C.sd;
      wave_sync();
      put<R>(s.tile, L, M);
      if (L.h == 0) {
        s.v0[j] = b_up;
        s.v1[j] = b_dn;
      }
      wave_sync();
      Mat<R> X;
      if (inverse_ns<R, false>(M, s.tile, s.tile2, X, L)) {  // M_up^-1
        Iup = left_solve_tile<R>(s.tile, b_up, L);            // or I_up by a solve
      } else {
        put<R>(s.tile, L, X);
        wave_sync();
        Iup = matvec<R>(s.tile, s.v0, L);
      }
      wave_sync();
      put<R>(s.tile, L, M2);
      wave_sync();
      if (inverse_ns<R, false>(M2, s.tile, s.tile2, X, L)) {  // M_dn^-1
        Idn = left_solve_tile<R>(s.tile, b_dn, L);             // or I_dn by a solve
      } else {
        put<R>(s.tile, L, X);
        wave_sync();
        Idn = matvec<R>(s.tile, s.v1, L);
      }
      wave_sync();
    } else {
      put<R>(s.tile, L, C.Tab);
      put<R>(s.tile2, L, C.Rba);
      if (L.h == 0) {
        s.v0[j] = Itd;
        s.v1[j] = Ibu;
      }
      wave_sync();
      const double tti = matvec<R>(s.tile, s.v0, L);
      const double rbi = matvec<R>(s.tile2, s.v1, L);
      wave_sync();
      Idn = tti + (rbi + C.sd);
      Iup = Ibu;
    }
    emit(l, Iup, Idn);
    if (l < Lay) add16_below(C, layer_ref<ADJ, F32>(a, lay, lflag, gw, l, mu_j, wt_j), l, L, s);
  }
}


}  // namespace w16
}  // namespace dev
}  // namespace adrt
