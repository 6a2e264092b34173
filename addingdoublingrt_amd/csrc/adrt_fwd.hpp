// adrt_fwd.hpp -- forward kernels of the batched adding-doubling solve (gfx950).
//
// Pipeline for one chunk of wavenumbers (all device-side, one stream):
//   prep_kernel    one thread per wavenumber: delta-M scaling of (tau, omega),
//                  cumulative scaled optical depth (same sequential order as
//                  src/solver.cpp:430-514 of the reference)
//   layer_kernel   one lane group per (wavenumber, layer): Legendre phase
//                  matrices + Hansen scaling, thin-layer init, doubling
//                  recursion; writes R, T and the combined (thermal + solar)
//                  layer sources to the layer store in HBM
//   adding_kernel  one lane group per wavenumber: bottom-up composite sweep
//                  (stores R_ab, T_ba, s_up of every base composite), TOA
//                  fluxes, then the top-down sweep with the interface solves
//                  and the flux / mean-intensity / divergence epilogue
//
// Thermal and solar sources travel as one combined pair through the adding
// sweeps: every adding formula is linear in the sources with the same
// operators (src/adding.hpp:31-62), so the sum equals the reference's separate
// propagation followed by s + s_solar at the interfaces.
#pragma once

#include "adrt_device.hpp"

namespace adrt {
namespace dev {

struct FwdArgs {
  // shape of this chunk
  int W;        // wavenumbers in the chunk
  int w0;       // first wavenumber of the chunk in the caller's arrays
  int L;        // layers
  int M;        // moments per layer
  int n_real;   // quadrature order requested (<= padded N)
  int shared_moments, delta_m, diffusion, rev, has_solar;
  double albedo, surf_em, top_em, F0, mu0, scaling, xfac;
  // caller inputs
  const double* tau;
  const double* ssa;
  const double* mom;
  const double* B;       // [W][L+1] or null
  const double* surf_w;  // [W] or null
  const double* top_w;   // [W] or null
  // caller outputs
  double* fu;
  double* fd;
  double* mi;
  double* fdir;
  double* fdiv;
  // quadrature tables (padded to N: w = 0, mu = 1, P_l(mu) = 0 on pad nodes)
  const double* mu;   // [N]
  const double* wt;   // [N]
  const double* Pl;   // [Lmax][N]
  const double* Pl0;  // [Lmax] Legendre polynomials at mu0
  const double* phase;  // [L][PH16] phase sums per layer when the moments are shared (N = 16), else null
  // workspace (chunk-local)
  double* tau_s;    // [W][L]   scaled tau used by doubling
  double* omega_s;  // [W][L]   scaled omega
  double* tau_cum;  // [W][L+1] exclusive prefix of tau_s (index L = column total)
  double* fdm;      // [W][L]   delta-M truncation f (mode 3)
  int* mode;        // [W][L]   0 no phase, 1 all M moments, 2 first min(M,2n), 3 delta-M
  double* lay;      // [W][L][2N^2 + 2N]   R, T (column-major), s_up, s_down
  int* lflag;       // [W][L]   is_scattering
  int* redo;        // [W][L]   mixed form: layer recomputed by the FP64 kernel
  int* pord;        // [W][L]   pair mode (orders <= 8 at 16): each column's layers by
                    //          doubling count, non-scattering first, pairs (2q, 2q + 1)
                    //          (pair_order_kernel); null = off
  int* pflag;       // [W][L]   pair mode: the layer runs in a pair (both layers of it scatter)
  int* ucol;        // [W]      N = 16 adding in the reference's composite form (consecutive
                    //          scattering layers with unstable doubling starts)
  int* ulist;       // [W]      the flagged columns, in no set order, so that the composite
                    //          kernels' first waves take them all (null: each wave checks
                    //          ucol of its own column)
  int* ucount;      // [1]      entries of ulist (zeroed before the prep kernel)
  double* rb;       // [W][L][2N^2 + 4N]   adding records (base composites; N = 16: the
                    //                     downward-sweep operators of each layer)
  double* basis;    // [W][L][2N] thermal source basis or null: y, z of the layer
                    // (s_up = y B_bar + z B_d, s_dn = y B_bar - z B_d), element j
  // reverse pass (N = 16 kernels)
  double* Iv;         // [W][L+1][2][N] forward interface intensities (up, down); null = not kept
  double* Av;         // [W][L+1][2][N] adjoint interface vectors (up, down)
  const double* cfu;  // [W][L+1] cotangents of the outputs in the caller's order; null = zero
  const double* cfd;
  const double* cJ;
  const double* cdir;
  const double* cdiv;
};

__device__ __forceinline__ int lin(const FwdArgs& a, int l) { return a.rev ? a.L - 1 - l : l; }
__device__ __forceinline__ int lev(const FwdArgs& a, int l) { return a.rev ? a.L - l : l; }

__device__ __forceinline__ int ipow0(double omega) {  // src/doubling.hpp:20-26
  return omega < 0.01 ? 4 : (omega < 0.1 ? 10 : 16);
}
__device__ __forceinline__ int doubling_count(double tau, double omega_in) {  // src/doubling.hpp:108-112
  const double omega = omega_in > 1.0 ? 1.0 : omega_in;
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(omega);
  return nn0 < 1 ? 1 : nn0;
}

// A scattering layer whose doubling start tau0 = tau / 2^nn lies beyond the
// smallest quadrature cosine mu_min (reachable only at omega < 0.01, where
// ipow0 = 4): the thin-layer transmission 1 - tau0 / mu has negative entries
// and the doubling amplifies them by orders of magnitude before they cancel.
__device__ __forceinline__ bool unstable_start(double tau_s, double omega_s, double mu_min) {
  if (!(tau_s > 0.0 && omega_s > 0.0)) return false;
  const double om = omega_s > 1.0 ? 1.0 : omega_s;
  const int nn0 = (int)(log(tau_s) / log(2.0)) + ipow0(om);
  const int nn = nn0 < 1 ? 1 : nn0;
  return tau_s * ldexp(1.0, -nn) > mu_min;
}

// Thick, strongly scattering layers (omega >= 0.8, tau >= 4): their FP32
// reverse doubling loses up to a few 1e-4 of the omega gradient (measured on
// the 65536 x 100 and 16384 x 100 N = 16 workloads: the columns beyond 1e-4
// each had such a layer, tau 7 - 10, omega 0.88 - 0.999).
__device__ __forceinline__ bool thick_conservative(double tau_s, double omega_s) {
  return omega_s >= 0.8 && tau_s >= 4.0;
}

// ---------------------------------------------------------------------------
// prep: delta-M and cumulative optical depth (reference src/solver.cpp:430-513),
// and the scattering layers the reduced-precision forms leave to FP64 (redo):
// those with an unstable doubling start and their neighbours in the column
// (whose rounding errors the unstable layers amplify), and thick, strongly
// scattering layers
// -- and the columns whose N = 16 adding runs in the reference's composite
// form (two consecutive scattering layers with unstable starts)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) prep_kernel(FwdArgs a) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.W) return;
  const long gw = (long)(a.w0 + w);
  const int L = a.L, M = a.M, twoM = 2 * a.n_real;
  const double mu_min = a.mu[0];
  double cum = 0.0;
  bool unst_prev = false;  // layer l - 1 has an unstable start
  bool unst_scat = false;  // the last scattering layer above has an unstable start
  int ucol = 0;
  for (int l = 0; l < L; ++l) {
    const int li = lin(a, l);
    const double tau = a.tau[gw * L + li];
    const double om = a.ssa[gw * L + li];
    double tl = tau, ol = om, f = 0.0;
    int mode = 0;
    if (om > 0.0 && tau > 0.0) {
      if (a.delta_m) {
        const double* chi = a.shared_moments ? a.mom + (long)li * M : a.mom + (gw * L + li) * M;
        f = (M > twoM) ? chi[twoM] : 0.0;
        if (f > 1e-12 && f < 1.0 - 1e-12) {
          const double omf = om * f;
          tl = (1.0 - omf) * tau;
          ol = om * (1.0 - f) / (1.0 - omf);
          mode = 3;
        } else {
          mode = 2;
        }
      } else {
        mode = 1;
      }
    }
    const long t = (long)w * L + l;
    a.tau_s[t] = tl;
    a.omega_s[t] = ol;
    a.fdm[t] = f;
    a.mode[t] = mode;
    a.tau_cum[(long)w * (L + 1) + l] = cum;
    cum += tl;
    const bool unst = mode != 0 && unstable_start(tl, ol, mu_min);
    a.redo[t] = (mode != 0 && (unst || thick_conservative(tl, ol))) ? 1 : 0;
    if (unst || unst_prev) {  // the scattering neighbours of an unstable layer
      if (l > 0 && a.mode[t - 1] != 0) a.redo[t - 1] = 1;
      if (mode != 0) a.redo[t] = 1;
    }
    unst_prev = unst;
    if (mode != 0) {
      if (unst && unst_scat) ucol = 1;
      unst_scat = unst;
    }
  }
  a.tau_cum[(long)w * (L + 1) + L] = cum;
  if (a.ucol) a.ucol[w] = ucol;
  if (ucol && a.ulist) a.ulist[atomicAdd(a.ucount, 1)] = w;
}

// The column a wave (or lane group) of a composite-form adding kernel takes:
// entry w of the prep kernel's list of flagged columns, so that the first
// waves of the grid -- dispatched first, beside the base sweep -- hold them
// all; without a list, column w itself if it is flagged.  Returns false for a
// wave with no column; w is wave-uniform either way.
__device__ __forceinline__ bool composite_column(const FwdArgs& a, int& w) {
  if (a.ulist) {
    if (w >= *a.ucount) return false;
    w = a.ulist[w];
    return true;
  }
  return w < a.W && a.ucol && a.ucol[w];
}

template <int N>
struct Lds {
  // per-group LDS window in doubles; the stride is 2 (mod 16) doubles so that
  // the groups of one wave start on different LDS banks
  static constexpr int raw = N * N + 7 * N;
  static constexpr int stride = raw + ((18 - (raw % 16)) % 16);
  static constexpr int tile = 0;
  static constexpr int v0 = N * N, v1 = v0 + N, v2 = v1 + N, v3 = v2 + N;
  static constexpr int prow = v3 + N;        // 2N
  static constexpr int pval = prow + 2 * N;  // N
};

// ---------------------------------------------------------------------------
// layer: phase matrices + doubling (src/phase_matrix.hpp:61-181,
// src/doubling.hpp:29-230); one lane group per (wavenumber, layer)
// ---------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256) layer_kernel(FwdArgs a) {
  using R = double;
  constexpr int RP = Geo<N>::RP, G = Geo<N>::G, GPB = Geo<N>::GPB, H = Geo<N>::H;
  constexpr int S = Lds<N>::stride;
  __shared__ R smem[GPB * S];
  const int g = threadIdx.x / G;
  const int j = (threadIdx.x % G) / H;
  const int h = threadIdx.x % H;
  const long task = (long)blockIdx.x * GPB + g;
  const int L = a.L;
  if (task >= (long)a.W * L) return;
  const int w = (int)(task / L), l = (int)(task % L);
  const long gw = (long)(a.w0 + w);
  R* sm = smem + g * S;
  R* tile = sm + Lds<N>::tile;

  constexpr int REC = 2 * N * N + 2 * N;
  R* out = a.lay + task * REC;

  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  const int mode = a.mode[task];
  const double Btop = a.B ? a.B[gw * (L + 1) + lev(a, l)] : 0.0;
  const double Bbot = a.B ? a.B[gw * (L + 1) + lev(a, l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (tau > 0.0) ? (Bbot - Btop) / tau : 0.0;
  const double mu_j = a.mu[j], wt_j = a.wt[j];

  if (tau <= 0.0) {  // transparent layer (src/doubling.hpp:52-53)
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      out[j * N + h * RP + r] = 0.0;
      out[N * N + j * N + h * RP + r] = (h * RP + r == j) ? 1.0 : 0.0;
    }
    if (h == 0) {
      out[2 * N * N + j] = 0.0;
      out[2 * N * N + N + j] = 0.0;
      if (a.basis) {
        a.basis[task * 2 * N + j] = 0.0;
        a.basis[task * 2 * N + N + j] = 0.0;
      }
    }
    if (j == 0 && h == 0) a.lflag[task] = 0;
    return;
  }
  if (omega_in <= 0.0) {  // pure absorption (src/doubling.hpp:56-77)
    const double tex = -tau / mu_j;
    const double trans = (tex > -200.0) ? exp(tex) : 0.0;
    const double omt = 1.0 - trans;
    const double slope = mu_j * omt - 0.5 * tau * (1.0 + trans);
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      out[j * N + h * RP + r] = 0.0;
      out[N * N + j * N + h * RP + r] = (h * RP + r == j) ? trans : 0.0;
    }
    if (h == 0) {
      out[2 * N * N + j] = B_bar * omt + B_d * slope;
      out[2 * N * N + N + j] = B_bar * omt - B_d * slope;
      if (a.basis) {
        a.basis[task * 2 * N + j] = omt;
        a.basis[task * 2 * N + N + j] = slope;
      }
    }
    if (j == 0 && h == 0) a.lflag[task] = 0;
    return;
  }

  // ---- scattering layer ----
  const double omega = omega_in > 1.0 ? 1.0 : omega_in;
  const double con = 2.0 * omega * kPi;
  const int li = lin(a, l);
  const double* chi = a.shared_moments ? a.mom + (long)li * a.M : a.mom + (gw * L + li) * a.M;
  const int twoM = 2 * a.n_real;
  int Lm = a.M;
  double f = 0.0, inv1mf = 1.0;
  if (mode == 3) {
    f = a.fdm[task];
    inv1mf = 1.0 / (1.0 - f);
    Lm = twoM;
  } else if (mode == 2) {
    Lm = a.M < twoM ? a.M : twoM;
  }

  // phase matrices, rows h*RP.. of column j (src/phase_matrix.hpp:73-91)
  R Ppp[RP], Ppm[RP];
#pragma unroll
  for (int r = 0; r < RP; ++r) { Ppp[r] = 0.0; Ppm[r] = 0.0; }
  double sign = 1.0;
  double spj = 0.0, smj = 0.0;  // solar phase vectors, element j
  const double* Ph = a.Pl + h * RP;
  for (int ll = 0; ll < Lm; ++ll) {
    const double chil = (mode == 3) ? (chi[ll] - f) * inv1mf : chi[ll];
    const double cl = (2 * ll + 1) * chil;
    const double plj = a.Pl[ll * N + j];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const double term = cl * Ph[ll * N + r] * plj;
      Ppp[r] += term;
      Ppm[r] += sign * term;
    }
    if (a.has_solar) {
      const double term = cl * plj * a.Pl0[ll];
      spj += term;
      smj += sign * term;
    }
    sign = -sign;
  }
  {
    // Hansen column normalisation (src/phase_matrix.hpp:93-110), summed over
    // rows in order 0..N-1 across the two halves of the column
    double csum = 0.0;
    if constexpr (H == 2) {
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        Ppp[r] = Ppp[r] / (2.0 * kPi);
        Ppm[r] = Ppm[r] / (2.0 * kPi);
      }
      double part = 0.0;
      if (h == 0)
        for (int r = 0; r < RP; ++r) part += (Ppp[r] + Ppm[r]) * a.wt[r];
      const double lower = swap_pair(part);
      if (h == 1) {
        part = lower;
        for (int r = 0; r < RP; ++r) part += (Ppp[r] + Ppm[r]) * a.wt[RP + r];
      }
      const double upper = swap_pair(part);
      csum = (h == 1) ? part : upper;
    } else {
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        Ppp[r] = Ppp[r] / (2.0 * kPi);
        Ppm[r] = Ppm[r] / (2.0 * kPi);
        csum += (Ppp[r] + Ppm[r]) * a.wt[r];
      }
    }
    if (csum > 0.0) {
      const double corr = 1.0 / (2.0 * kPi * csum);
#pragma unroll
      for (int r = 0; r < RP; ++r) { Ppp[r] *= corr; Ppm[r] *= corr; }
    }
  }
  double pplus = 0.0, pminus = 0.0;
  if (a.has_solar) {  // src/phase_matrix.hpp:153-180
    pplus = spj / (2.0 * kPi);
    pminus = smj / (2.0 * kPi);
    const double s = group_sum_ordered<N, R>((pplus + pminus) * wt_j, j, h, sm + Lds<N>::pval);
    if (s > 0.0) {
      const double corr = 1.0 / (2.0 * kPi * s);
      pplus *= corr;
      pminus *= corr;
    }
  }

  // thin-layer initialisation (src/doubling.hpp:83-147)
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(omega);
  const int nn = nn0 < 1 ? 1 : nn0;
  const double tau0 = tau * ldexp(1.0, -nn);
  R Rm[RP], Tm[RP];
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int i = h * RP + r;
    const double mui = a.mu[i];
    const double gpp = (((i == j) ? 1.0 : 0.0) + (-con) * (Ppp[r] * wt_j)) / mui;
    const double gpm = con * (Ppm[r] * wt_j) / mui;
    Tm[r] = ((i == j) ? 1.0 : 0.0) - tau0 * gpp;
    Rm[r] = tau0 * gpm;
  }
  double y = (1.0 - omega) * tau0 / mu_j;
  double z = 0.0;
  double sus = 0.0, sds = 0.0, gam = 0.0;
  if (a.has_solar) {
    const double Ftop = a.F0 * exp(-a.tau_cum[(long)w * (L + 1) + l] / a.mu0);
    const double base = omega * tau0 / mu_j * Ftop;
    sus = base * pminus;
    sds = base * pplus;
    gam = exp(-tau0 / a.mu0);
  }
  double gk = 0.5 * tau0;

  R* v0 = sm + Lds<N>::v0;
  R* v1 = sm + Lds<N>::v1;
  R* v2 = sm + Lds<N>::v2;
  R* v3 = sm + Lds<N>::v3;
  R* prow = sm + Lds<N>::prow;
  R* pval = sm + Lds<N>::pval;

  // doubling recursion (src/doubling.hpp:149-209)
  for (int k = 0; k < nn; ++k) {
    R A[RP], X[RP], Y[RP];
    put_col<N, R>(tile, j, h, Rm);
    if (h == 0) {
      v0[j] = sds;
      v1[j] = sus;
    }
    wave_sync();
    eye_minus_mul<N, R>(tile, Rm, j, h, A);  // I - R^2
    double b1 = 0.0, b2 = 0.0;
    if (a.has_solar) {
      double bs[2];
      matvecs<N, 2, R>(tile, v0, j, bs);  // R s_down_sol, R s_up_sol
      b1 = bs[0];
      b2 = bs[1];
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < RP; ++r) X[r] = Tm[r];
    right_solve<N, R, false>(A, X, j, h, prow, pval, tile);  // TG = T (I - R^2)^-1
    put_col<N, R>(tile, j, h, X);
    const double zpgy = z + gk * y;
    if (h == 0) {
      v0[j] = zpgy;
      v1[j] = y;
      v2[j] = b1 + gam * sus;
      v3[j] = gam * b2 + sds;
    }
    wave_sync();
    mul<N, R>(tile, Rm, h, Y);  // TGR = TG R
    mul<N, R>(tile, Tm, h, A);  // T_new = TG T
    double a1, a3, c1 = 0.0, c2 = 0.0;
    if (a.has_solar) {
      double s4[4];
      matvecs<N, 4, R>(tile, v0, j, s4);
      a1 = s4[0];
      a3 = s4[1];
      c1 = s4[2];
      c2 = s4[3];
    } else {
      double s2[2];
      matvecs<N, 2, R>(tile, v0, j, s2);
      a1 = s2[0];
      a3 = s2[1];
    }
    wave_sync();
    put_col<N, R>(tile, j, h, Y);
    wave_sync();
    mul<N, R>(tile, Tm, h, X);  // TGR T
#pragma unroll
    for (int r = 0; r < RP; ++r) Rm[r] = Rm[r] + X[r];  // R_new = R + TGR T
    double s2[2];
    matvecs<N, 2, R>(tile, v0, j, s2);
    const double a2 = s2[0], a4 = s2[1];
    wave_sync();
#pragma unroll
    for (int r = 0; r < RP; ++r) Tm[r] = A[r];
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

  // layer assembly (src/doubling.hpp:211-229), sources combined
  store_col<N, R>(out, j, h, Rm);
  store_col<N, R>(out + N * N, j, h, Tm);
  if (h == 0) {
    out[2 * N * N + j] = (y * B_bar + z * B_d) + sus;
    out[2 * N * N + N + j] = (y * B_bar - z * B_d) + sds;
    if (a.basis) {
      a.basis[task * 2 * N + j] = y;
      a.basis[task * 2 * N + N + j] = z;
    }
  }
  if (j == 0 && h == 0) a.lflag[task] = 1;
}

// ---------------------------------------------------------------------------
// adding sweeps + interfaces (src/adding.hpp:69-258, src/solver.cpp:516-700)
// ---------------------------------------------------------------------------

// running composite: slices of R_ab, R_ba, T_ab, T_ba; s_up, s_down element j
template <int N>
struct Comp {
  double Rab[Geo<N>::RP], Rba[Geo<N>::RP], Tab[Geo<N>::RP], Tba[Geo<N>::RP];
  double su, sd;
  int scat;
};

// (X v)_j for X stored column-major in global memory, v in an LDS slot
template <int N>
__device__ __forceinline__ double gmatvec(const double* X, const double* v, int j) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) s = fma_(X[k * N + j], v[k], s);
  return s;
}

// rbase step: new = add(top = stored layer, bot = C)  (src/adding.hpp:246-258)
template <int N>
__device__ void add_layer_on_top(Comp<N>& C, const double* rec, int lscat, int j, int h,
                                 double* sm) {
  using R = double;
  constexpr int RP = Geo<N>::RP;
  R* tile = sm + Lds<N>::tile;
  R* v0 = sm + Lds<N>::v0;
  R* v1 = sm + Lds<N>::v1;
  R* v2 = sm + Lds<N>::v2;
  R* v3 = sm + Lds<N>::v3;
  R* prow = sm + Lds<N>::prow;
  R* pval = sm + Lds<N>::pval;
  const R* Rt = rec;
  const R* Tt = rec + N * N;
  const R tsu = rec[2 * N * N + j];
  const R tsd = rec[2 * N * N + N + j];
  R A[RP], X[RP], Lc[RP];
  if (!lscat) {
    // NonscatTop (src/adding.hpp:117-174): t = diag(top.T_ab)
    const R tj = Tt[j * N + j];
    if (h == 0) {
      v0[j] = tj;
      v1[j] = tsd;
    }
    put_col<N, R>(tile, j, h, C.Rab);
    wave_sync();
    const R rbc = matvec<N, R>(tile, v1, j);  // bot.R_ab s1_down
    wave_sync();
    put_col<N, R>(tile, j, h, C.Tab);
    wave_sync();
    const R tbc = matvec<N, R>(tile, v1, j);  // bot.T_ab s1_down
    R tv[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) tv[r] = v0[h * RP + r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      C.Rab[r] = tv[r] * C.Rab[r] * tj;
      C.Tab[r] = C.Tab[r] * tj;
      C.Tba[r] = tv[r] * C.Tba[r];
    }
    const R nsu = tsu + tj * (C.su + rbc);
    const R nsd = C.sd + tbc;
    C.su = nsu;
    C.sd = nsd;
    return;  // scat stays bot's
  }
  if (!C.scat) {
    // NonscatBot (src/adding.hpp:181-239): t = diag(bot.T_ab)
    const R tj = diag_of<N, R>(C.Tab, j, h);
    if (h == 0) {
      v0[j] = tj;
      v1[j] = C.su;
    }
    wave_sync();
    R tv[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) tv[r] = v0[h * RP + r];
    const R tba = gmatvec<N>(Tt, v1, j);  // top.T_ba s2_up
    const R rba = gmatvec<N>(Rt, v1, j);  // top.R_ba s2_up
    wave_sync();
    load_col<N, R>(Rt, j, h, Lc);
#pragma unroll
    for (int r = 0; r < RP; ++r) { C.Rab[r] = Lc[r]; C.Rba[r] = tv[r] * Lc[r] * tj; }
    load_col<N, R>(Tt, j, h, Lc);
#pragma unroll
    for (int r = 0; r < RP; ++r) { C.Tab[r] = tv[r] * Lc[r]; C.Tba[r] = Lc[r] * tj; }
    const R nsu = tsu + tba;
    const R nsd = C.sd + tj * (tsd + rba);
    C.su = nsu;
    C.sd = nsd;
    C.scat = 1;
    return;
  }
  // General (src/adding.hpp:69-110)
  put_col<N, R>(tile, j, h, C.Rab);
  if (h == 0) v0[j] = tsd;
  wave_sync();
  load_col<N, R>(Rt, j, h, Lc);
  eye_minus_mul<N, R>(tile, Lc, j, h, A);             // A1 = I - bot.R_ab top.R_ba
  const R rhs_up = C.su + matvec<N, R>(tile, v0, j);  // s2_up + R_bc s1_down
  wave_sync();
  load_col<N, R>(Tt, j, h, X);
  right_solve<N, R>(A, X, j, h, prow, pval, tile);    // D1 = top.T_ba A1^-1
  put_col<N, R>(tile, j, h, Lc);
  if (h == 0) v1[j] = C.su;
  wave_sync();
  eye_minus_mul<N, R>(tile, C.Rab, j, h, A);          // A2 = I - top.R_ba bot.R_ab
  const R rhs_dn = tsd + matvec<N, R>(tile, v1, j);   // s1_down + R_ba s2_up
  wave_sync();
  right_solve<N, R>(A, C.Tab, j, h, prow, pval, tile);  // D2 = bot.T_ab A2^-1
  if (h == 0) {
    v2[j] = rhs_up;
    v3[j] = rhs_dn;
  }
  put_col<N, R>(tile, j, h, C.Tab);  // D2
  wave_sync();
  mul<N, R>(tile, Lc, h, A);                          // D2 top.R_ba
  const R nsd = C.sd + matvec<N, R>(tile, v3, j);
  load_col<N, R>(Tt, j, h, Lc);
  mul<N, R>(tile, Lc, h, C.Tab);                      // T_ab = D2 top.T_ab
  wave_sync();
  put_col<N, R>(tile, j, h, A);
  wave_sync();
  mul_acc<N, R>(tile, C.Tba, h, C.Rba);               // R_ba = bot.R_ba + D2 R_ba T_ba
  wave_sync();
  put_col<N, R>(tile, j, h, X);                       // D1
  wave_sync();
  mul<N, R>(tile, C.Rab, h, A);                       // D1 bot.R_ab
  const R nsu = tsu + matvec<N, R>(tile, v2, j);
  mul<N, R>(tile, C.Tba, h, X);                       // T_ba = D1 bot.T_ba
  wave_sync();
#pragma unroll
  for (int r = 0; r < RP; ++r) C.Tba[r] = X[r];
  put_col<N, R>(tile, j, h, A);
  wave_sync();
  load_col<N, R>(Rt, j, h, X);
  mul<N, R>(tile, Lc, h, A);                          // (D1 R_ab) top.T_ab
#pragma unroll
  for (int r = 0; r < RP; ++r) C.Rab[r] = X[r] + A[r];  // R_ab = top.R_ab + ...
  wave_sync();
  C.su = nsu;
  C.sd = nsd;
  C.scat = 1;
}

// rtop step: new = add(top = C, bot = stored layer)
template <int N>
__device__ void add_layer_below(Comp<N>& C, const double* rec, int lscat, int j, int h,
                                double* sm) {
  using R = double;
  constexpr int RP = Geo<N>::RP;
  R* tile = sm + Lds<N>::tile;
  R* v0 = sm + Lds<N>::v0;
  R* v1 = sm + Lds<N>::v1;
  R* v2 = sm + Lds<N>::v2;
  R* v3 = sm + Lds<N>::v3;
  R* prow = sm + Lds<N>::prow;
  R* pval = sm + Lds<N>::pval;
  const R* Rb = rec;
  const R* Tb = rec + N * N;
  const R bsu = rec[2 * N * N + j];
  const R bsd = rec[2 * N * N + N + j];
  R A[RP], X[RP], Lc[RP];
  if (!C.scat) {
    // NonscatTop with top = C: t = diag(C.T_ab)
    const R tj = diag_of<N, R>(C.Tab, j, h);
    if (h == 0) {
      v0[j] = tj;
      v1[j] = C.sd;
    }
    wave_sync();
    R tv[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) tv[r] = v0[h * RP + r];
    const R rbc = gmatvec<N>(Rb, v1, j);  // bot.R_ab s1_down
    const R tbc = gmatvec<N>(Tb, v1, j);  // bot.T_ab s1_down
    wave_sync();
    load_col<N, R>(Rb, j, h, Lc);
#pragma unroll
    for (int r = 0; r < RP; ++r) { C.Rab[r] = tv[r] * Lc[r] * tj; C.Rba[r] = Lc[r]; }
    load_col<N, R>(Tb, j, h, Lc);
#pragma unroll
    for (int r = 0; r < RP; ++r) { C.Tab[r] = Lc[r] * tj; C.Tba[r] = tv[r] * Lc[r]; }
    const R nsu = C.su + tj * (bsu + rbc);
    const R nsd = bsd + tbc;
    C.su = nsu;
    C.sd = nsd;
    C.scat = lscat;
    return;
  }
  if (!lscat) {
    // NonscatBot with bot = layer: t = diag(layer T)
    const R tj = Tb[j * N + j];
    if (h == 0) {
      v0[j] = tj;
      v1[j] = bsu;
    }
    put_col<N, R>(tile, j, h, C.Tba);
    wave_sync();
    const R tba = matvec<N, R>(tile, v1, j);  // top.T_ba s2_up
    wave_sync();
    put_col<N, R>(tile, j, h, C.Rba);
    wave_sync();
    const R rba = matvec<N, R>(tile, v1, j);  // top.R_ba s2_up
    R tv[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) tv[r] = v0[h * RP + r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      C.Rba[r] = tv[r] * C.Rba[r] * tj;
      C.Tab[r] = tv[r] * C.Tab[r];
      C.Tba[r] = C.Tba[r] * tj;
    }
    const R nsu = C.su + tba;
    const R nsd = bsd + tj * (C.sd + rba);
    C.su = nsu;
    C.sd = nsd;
    return;
  }
  // General with top = C, bot = layer
  load_col<N, R>(Rb, j, h, Lc);
  put_col<N, R>(tile, j, h, Lc);
  if (h == 0) v0[j] = C.sd;
  wave_sync();
  eye_minus_mul<N, R>(tile, C.Rba, j, h, A);          // A1 = I - bot.R_ab top.R_ba
  const R rhs_up = bsu + matvec<N, R>(tile, v0, j);   // s2_up + R_bc s1_down
  wave_sync();
  right_solve<N, R>(A, C.Tba, j, h, prow, pval, tile);  // D1 = top.T_ba A1^-1
  put_col<N, R>(tile, j, h, C.Rba);
  if (h == 0) v1[j] = bsu;
  wave_sync();
  eye_minus_mul<N, R>(tile, Lc, j, h, A);             // A2 = I - top.R_ba bot.R_ab
  const R rhs_dn = C.sd + matvec<N, R>(tile, v1, j);  // s1_down + R_ba s2_up
  wave_sync();
  load_col<N, R>(Tb, j, h, X);
  right_solve<N, R>(A, X, j, h, prow, pval, tile);    // D2 = bot.T_ab A2^-1
  if (h == 0) {
    v2[j] = rhs_up;
    v3[j] = rhs_dn;
  }
  put_col<N, R>(tile, j, h, C.Tba);                   // D1
  wave_sync();
  mul<N, R>(tile, Lc, h, A);                          // D1 bot.R_ab
  const R nsu = C.su + matvec<N, R>(tile, v2, j);
  load_col<N, R>(Tb, j, h, Lc);
  mul<N, R>(tile, Lc, h, C.Tba);                      // T_ba = D1 bot.T_ba
  wave_sync();
  put_col<N, R>(tile, j, h, A);
  wave_sync();
  mul_acc<N, R>(tile, C.Tab, h, C.Rab);               // R_ab = top.R_ab + D1 R_ab T_ab
  wave_sync();
  put_col<N, R>(tile, j, h, X);                       // D2
  wave_sync();
  mul<N, R>(tile, C.Rba, h, A);                       // D2 top.R_ba
  const R nsd = bsd + matvec<N, R>(tile, v3, j);
  mul<N, R>(tile, C.Tab, h, X);                       // T_ab = D2 top.T_ab
  wave_sync();
#pragma unroll
  for (int r = 0; r < RP; ++r) C.Tab[r] = X[r];
  put_col<N, R>(tile, j, h, A);
  wave_sync();
  load_col<N, R>(Rb, j, h, X);
  mul<N, R>(tile, Lc, h, A);                          // (D2 R_ba) bot.T_ba
#pragma unroll
  for (int r = 0; r < RP; ++r) C.Rba[r] = X[r] + A[r];  // R_ba = bot.R_ba + ...
  wave_sync();
  C.su = nsu;
  C.sd = nsd;
  C.scat = 1;
}

// x = M^-1 b for M given by column slices, b distributed (element j).
template <int N>
__device__ __forceinline__ double left_solve(const double (&Mc)[Geo<N>::RP], double b, int j,
                                             int h, double* sm) {
  constexpr int RP = Geo<N>::RP;
  double* tile = sm + Lds<N>::tile;
  put_col<N, double>(tile, j, h, Mc);
  wave_sync();
  double a[RP];
  get_row<N, double>(tile, j, h, a);  // slice of column j of M^T
  wave_sync();
  return right_solve_vec<N, double>(a, b, j, h, sm + Lds<N>::prow, sm + Lds<N>::pval, tile);
}

// The reference's form, for the columns the prep kernel flags (consecutive
// scattering layers with unstable doubling starts; the N = 16 adding's
// composite kernel gives the reason); adding_kernel below does the others.
template <int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
adding_composite_kernel(FwdArgs a) {
  using R = double;
  constexpr int RP = Geo<N>::RP, G = Geo<N>::G, GPB = Geo<N>::GPB, H = Geo<N>::H;
  constexpr int S = Lds<N>::stride;
  __shared__ R smem[GPB * S];
  const int g = threadIdx.x / G;
  const int j = (threadIdx.x % G) / H;
  const int h = threadIdx.x % H;
  int w = blockIdx.x * GPB + g;
  if (!composite_column(a, w)) return;
  const long gw = (long)(a.w0 + w);
  const int L = a.L;
  R* sm = smem + g * S;
  R* tile = sm + Lds<N>::tile;
  R* v0 = sm + Lds<N>::v0;
  R* v1 = sm + Lds<N>::v1;
  constexpr int REC = 2 * N * N + 2 * N;
  constexpr int RBR = 2 * N * N + N;
  const double* lay = a.lay + (long)w * L * REC;
  const int* lflag = a.lflag + (long)w * L;
  // the column's region of the adding store is L (2N^2 + 4N) doubles, as the
  // base sweep's (RbRec<N>::size per layer): this kernel runs beside
  // adding_kernel, so its columns must not reach into that kernel's columns
  double* rbw = a.rb + (long)w * L * (2 * N * N + 4 * N);
  const double mu_j = a.mu[j], wt_j = a.wt[j];
  const double* Bw = a.B ? a.B + gw * (L + 1) : nullptr;
  auto Blev = [&](int l) -> double { return Bw ? Bw[lev(a, l)] : 0.0; };

  // boundary conditions (src/solver.cpp:385-419, 516-551, 585-603)
  const double Bs = a.surf_w ? a.surf_w[gw] : a.surf_em;
  const double Btop = a.top_w ? a.top_w[gw] : a.top_em;
  const double A = a.albedo;
  const int has_surface = !a.diffusion && (A > 0.0 || Bs > 0.0);
  const int ltot = L + (has_surface ? 1 : 0);
  const double tau_tot = a.tau_cum[(long)w * (L + 1) + L];
  const double Itd = Btop;
  double Ibu = 0.0;
  if (a.diffusion) {
    const double Bb = Blev(L);
    const double dt = a.tau_s[(long)w * L + L - 1];
    const double dBdt = (dt > 0.0) ? (Bb - Blev(L - 1)) / dt : 0.0;
    Ibu = Bb + mu_j * dBdt;
  } else if (!has_surface) {
    Ibu = Bs;
  }

  Comp<N> C;
  // ---- bottom-up sweep (src/solver.cpp:560-572) ----
  if (has_surface) {
    const double rr = 2.0 * A * mu_j * wt_j * a.xfac;
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      C.Rab[r] = (h * RP + r < a.n_real) ? rr : 0.0;
      C.Rba[r] = C.Rab[r];
      C.Tab[r] = 0.0;
      C.Tba[r] = 0.0;
    }
    double ssol = 0.0;
    if (a.has_solar && A > 0.0) ssol = (A / kPi) * a.F0 * a.mu0 * exp(-tau_tot / a.mu0);
    C.su = (1.0 - A) * Bs + ssol;
    C.sd = 0.0;
    C.scat = A > 0.0;
  } else {
    const double* rec = lay + (long)(L - 1) * REC;
    load_col<N, R>(rec, j, h, C.Rab);
    load_col<N, R>(rec + N * N, j, h, C.Tab);
#pragma unroll
    for (int r = 0; r < RP; ++r) { C.Rba[r] = C.Rab[r]; C.Tba[r] = C.Tab[r]; }
    C.su = rec[2 * N * N + j];
    C.sd = rec[2 * N * N + N + j];
    C.scat = lflag[L - 1];
  }
  for (int n = 1;; ++n) {  // C == rbase[n]
    if (n < ltot) {        // keep base composites n = 1..ltot-1 for the interfaces
      double* o = rbw + (long)(n - 1) * RBR;
      store_col<N, R>(o, j, h, C.Rab);
      store_col<N, R>(o + N * N, j, h, C.Tba);
      if (h == 0) o[2 * N * N + j] = C.su;
    }
    if (n == ltot) break;
    const int k = ltot - 1 - n;
    add_layer_on_top<N>(C, lay + (long)k * REC, lflag[k], j, h, sm);
  }

  auto emit = [&](int l, double Iup, double Idn) {
    const double c = 2.0 * kPi * wt_j * mu_j;
    const double fu = group_sum_ordered<N, R>(c * Iup, j, h, sm + Lds<N>::pval);
    const double fdn = group_sum_ordered<N, R>(c * Idn, j, h, sm + Lds<N>::pval);
    double J = group_sum_ordered<N, R>(0.5 * wt_j * (Iup + Idn), j, h, sm + Lds<N>::pval);
    if (j == 0 && h == 0) {
      const long o = gw * (L + 1) + lev(a, l);
      double fdir = 0.0;
      if (a.has_solar) {
        fdir = (l == 0) ? a.F0 * a.mu0
                        : a.F0 * a.mu0 * exp(-a.tau_cum[(long)w * (L + 1) + l] / a.mu0);
        J += fdir * (1.0 / (4.0 * kPi * a.mu0));
      }
      a.fu[o] = fu * a.scaling;
      a.fd[o] = fdn;
      a.mi[o] = J;
      a.fdir[o] = fdir;
      if (a.fdiv) {
        const int lyr = (l == 0) ? 0 : l - 1;
        const double om = a.ssa[gw * L + lin(a, lyr)];
        a.fdiv[o] = 4.0 * kPi * (1.0 - om) * (J - Blev(l));
      }
    }
  };

  // ---- TOA (src/solver.cpp:628-642) ----
  {
    put_col<N, R>(tile, j, h, C.Rab);
    if (h == 0) {
      v0[j] = Itd;
      v1[j] = Ibu;
    }
    wave_sync();
    const double rI = matvec<N, R>(tile, v0, j);
    wave_sync();
    put_col<N, R>(tile, j, h, C.Tba);
    wave_sync();
    const double tI = matvec<N, R>(tile, v1, j);
    wave_sync();
    emit(0, rI + (tI + C.su), Itd);
  }

  // ---- top-down sweep with interfaces (src/solver.cpp:574-583, 644-678) ----
  {
    const double* rec = lay;
    load_col<N, R>(rec, j, h, C.Rab);
    load_col<N, R>(rec + N * N, j, h, C.Tab);
#pragma unroll
    for (int r = 0; r < RP; ++r) { C.Rba[r] = C.Rab[r]; C.Tba[r] = C.Tab[r]; }
    C.su = rec[2 * N * N + j];
    C.sd = rec[2 * N * N + N + j];
    C.scat = lflag[0];
  }
  for (int l = 1; l <= L; ++l) {  // C == rtop[l]
    const int n_base = ltot - l;
    double Iup, Idn;
    if (n_base > 0) {
      const double* o = rbw + (long)(n_base - 1) * RBR;
      const double sbu = o[2 * N * N + j];
      R Mc[RP], Lc[RP];
      // top.T_ab I_top_down
      put_col<N, R>(tile, j, h, C.Tab);
      if (h == 0) v0[j] = Itd;
      wave_sync();
      const double tti = matvec<N, R>(tile, v0, j);
      wave_sync();
      // base.T_ba I_bot_up
      if (h == 0) v1[j] = Ibu;
      load_col<N, R>(o + N * N, j, h, Lc);
      put_col<N, R>(tile, j, h, Lc);
      wave_sync();
      const double tbi = matvec<N, R>(tile, v1, j);
      wave_sync();
      // M_up = I - base.R_ab top.R_ba ; terms 2 and 3 (src/solver.cpp:24-46)
      if (h == 0) {
        v0[j] = tti;
        v1[j] = C.sd;
      }
      load_col<N, R>(o, j, h, Lc);
      put_col<N, R>(tile, j, h, Lc);
      wave_sync();
      eye_minus_mul<N, R>(tile, C.Rba, j, h, Mc);
      const double t2 = matvec<N, R>(tile, v0, j);
      const double t3 = matvec<N, R>(tile, v1, j);
      wave_sync();
      Iup = left_solve<N>(Mc, tbi + t2 + t3 + sbu, j, h, sm);
      // M_dn = I - top.R_ba base.R_ab (src/solver.cpp:51-73)
      put_col<N, R>(tile, j, h, C.Rba);
      if (h == 0) {
        v0[j] = tbi;
        v1[j] = sbu;
      }
      wave_sync();
      eye_minus_mul<N, R>(tile, Lc, j, h, Mc);
      const double u2 = matvec<N, R>(tile, v0, j);
      const double u3 = matvec<N, R>(tile, v1, j);
      wave_sync();
      Idn = left_solve<N>(Mc, tti + u2 + u3 + C.sd, j, h, sm);
    } else {
      put_col<N, R>(tile, j, h, C.Tab);
      if (h == 0) {
        v0[j] = Itd;
        v1[j] = Ibu;
      }
      wave_sync();
      const double tti = matvec<N, R>(tile, v0, j);
      wave_sync();
      put_col<N, R>(tile, j, h, C.Rba);
      wave_sync();
      const double rbi = matvec<N, R>(tile, v1, j);
      wave_sync();
      Idn = tti + (rbi + C.sd);
      Iup = Ibu;
    }
    emit(l, Iup, Idn);
    if (l < L) add_layer_below<N>(C, lay + (long)l * REC, lflag[l], j, h, sm);
  }
}


// ---------------------------------------------------------------------------
// adding by the base sweep on column slices: the same system as the N = 16
// adding16_kernel solves (see adrt_fwd16.hpp) -- one sweep of the base
// composites Rb(l), sb(l) from the bottom, then a downward sweep of
// matrix-vector products for the interface intensities.  Per scattering
// layer two LU solves with I - R_l Rb(l+1) (for G_l and d_l) and three
// products, against six solves and about twenty products for the composites
// of both sweeps and the interface solves.
// ---------------------------------------------------------------------------
// per layer: G_l (column-major; the diagonal of T_l for a non-scattering
// layer), Rb(l+1) (column-major), d_l, sb(l+1), whether Rb(l+1) is non-zero
template <int N>
struct RbRec {
  static constexpr int G = 0, R = N * N, D = 2 * N * N, S = D + N, F = S + N;
  static constexpr int size = 2 * N * N + 4 * N;
  static_assert(F < size, "adding record fields past the record");
};

template <int N>
struct Base {
  double R[Geo<N>::RP];  // Rb, this lane's slice of column j (all zero unless scat)
  double s;              // sb, element j
  int scat;
};

// layer rec on top of the base B: stores at o what the downward sweep needs at
// this layer, then replaces B by the base above the layer
template <int N>
__device__ void base_add(Base<N>& B, const double* rec, int lscat, double* o, int j, int h,
                         double* sm) {
  using R = double;
  using F = RbRec<N>;
  constexpr int RP = Geo<N>::RP;
  R* tile = sm + Lds<N>::tile;
  R* v0 = sm + Lds<N>::v0;
  R* v1 = sm + Lds<N>::v1;
  R* prow = sm + Lds<N>::prow;
  R* pval = sm + Lds<N>::pval;
  const R* Rl = rec;
  const R* Tl = rec + N * N;
  const R su = rec[2 * N * N + j];
  const R sd = rec[2 * N * N + N + j];
  if (B.scat) store_col<N, R>(o + F::R, j, h, B.R);
  if (h == 0) {
    o[F::S + j] = B.s;
    if (j == 0) o[F::F] = B.scat ? 1.0 : 0.0;
  }
  if (!lscat) {
    // T_l = diag(t), R_l = 0: G = T_l, d = s_dn, Rb = T_l Rb T_l,
    // sb = s_up + T_l (Rb s_dn + sb)
    const R tj = Tl[j * N + j];
    R rsd = 0.0;
    if (B.scat) {
      put_col<N, R>(tile, j, h, B.R);
      if (h == 0) {
        v0[j] = tj;
        v1[j] = sd;
      }
      wave_sync();
      rsd = matvec<N, R>(tile, v1, j);
      R tv[RP];
#pragma unroll
      for (int r = 0; r < RP; ++r) tv[r] = v0[h * RP + r];
      wave_sync();
#pragma unroll
      for (int r = 0; r < RP; ++r) B.R[r] = tv[r] * B.R[r] * tj;
    }
    if (h == 0) {
      o[F::G + j] = tj;
      o[F::D + j] = sd;
    }
    B.s = su + tj * (rsd + B.s);
    return;
  }
  if (h == 0) v0[j] = B.s;
  wave_sync();
  const R v = gmatvec<N>(Rl, v0, j) + sd;  // R_l sb(l+1) + s_dn
  R Lc[RP], Mc[RP];
  if (!B.scat) {
    // nothing below reflects: G = T_l, d = R_l sb + s_dn, Rb = R_l, sb = s_up + T_l sb
    const R ts = gmatvec<N>(Tl, v0, j);
    wave_sync();
    load_col<N, R>(Tl, j, h, Lc);
    store_col<N, R>(o + F::G, j, h, Lc);
    load_col<N, R>(Rl, j, h, B.R);
    if (h == 0) o[F::D + j] = v;
    B.s = su + ts;
    B.scat = 1;
    return;
  }
  wave_sync();
  load_col<N, R>(Rl, j, h, Lc);
  put_col<N, R>(tile, j, h, Lc);
  wave_sync();
  eye_minus_mul<N, R>(tile, B.R, j, h, Mc);  // M = I - R_l Rb(l+1)
  wave_sync();
  const R d = left_solve<N>(Mc, v, j, h, sm);  // d = M^-1 (R_l sb + s_dn)
  // G = M^-1 T_l as the right solve G^T M^T = T_l^T: this lane holds its slice
  // of row j of M (column j of M^T) and of row j of T_l, and gets its slice of
  // row j of G
  put_col<N, R>(tile, j, h, Mc);
  wave_sync();
  R ra[RP], rt[RP], Gc[RP];
  get_row<N, R>(tile, j, h, ra);
  wave_sync();
#pragma unroll
  for (int r = 0; r < RP; ++r) rt[r] = Tl[(h * RP + r) * N + j];
  right_solve<N, R>(ra, rt, j, h, prow, pval, tile);
  put_col<N, R>(tile, j, h, rt);  // G^T, column-major
  wave_sync();
  get_row<N, R>(tile, j, h, Gc);  // this lane's slice of column j of G
  wave_sync();
  store_col<N, R>(o + F::G, j, h, Gc);
  if (h == 0) o[F::D + j] = d;
  // Rb(l) = R_l + T_l (Rb(l+1) G), sb(l) = s_up + T_l (Rb(l+1) d + sb(l+1))
  put_col<N, R>(tile, j, h, B.R);
  if (h == 0) v1[j] = d;
  wave_sync();
  mul<N, R>(tile, Gc, h, Mc);  // Rb(l+1) G
  const R rd = matvec<N, R>(tile, v1, j);
  wave_sync();
  load_col<N, R>(Tl, j, h, Lc);
  put_col<N, R>(tile, j, h, Lc);
  if (h == 0) v0[j] = rd + B.s;
  wave_sync();
  mul<N, R>(tile, Mc, h, Gc);  // T_l Rb(l+1) G
  const R ts = matvec<N, R>(tile, v0, j);
  wave_sync();
  load_col<N, R>(Rl, j, h, Lc);
#pragma unroll
  for (int r = 0; r < RP; ++r) B.R[r] = Lc[r] + Gc[r];
  B.s = su + ts;
}

template <int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) adding_kernel(FwdArgs a) {
  using R = double;
  using F = RbRec<N>;
  constexpr int RP = Geo<N>::RP, G = Geo<N>::G, GPB = Geo<N>::GPB, H = Geo<N>::H;
  constexpr int S = Lds<N>::stride;
  __shared__ R smem[GPB * S];
  const int g = threadIdx.x / G;
  const int j = (threadIdx.x % G) / H;
  const int h = threadIdx.x % H;
  const int w = blockIdx.x * GPB + g;
  if (w >= a.W || (a.ucol && a.ucol[w])) return;
  const long gw = (long)(a.w0 + w);
  const int L = a.L;
  R* sm = smem + g * S;
  R* tile = sm + Lds<N>::tile;
  R* v0 = sm + Lds<N>::v0;
  constexpr int REC = 2 * N * N + 2 * N;
  const double* lay = a.lay + (long)w * L * REC;
  const int* lflag = a.lflag + (long)w * L;
  double* rbw = a.rb + (long)w * L * F::size;
  const double mu_j = a.mu[j], wt_j = a.wt[j];
  const double* Bw = a.B ? a.B + gw * (L + 1) : nullptr;
  auto Blev = [&](int l) -> double { return Bw ? Bw[lev(a, l)] : 0.0; };

  // boundary conditions (src/solver.cpp:385-419, 516-551, 585-603)
  const double Bs = a.surf_w ? a.surf_w[gw] : a.surf_em;
  const double Btop = a.top_w ? a.top_w[gw] : a.top_em;
  const double A = a.albedo;
  const int has_surface = !a.diffusion && (A > 0.0 || Bs > 0.0);
  const double tau_tot = a.tau_cum[(long)w * (L + 1) + L];
  const double Itd = Btop;
  double Ibu = 0.0;
  if (a.diffusion) {
    const double Bb = Blev(L);
    const double dt = a.tau_s[(long)w * L + L - 1];
    const double dBdt = (dt > 0.0) ? (Bb - Blev(L - 1)) / dt : 0.0;
    Ibu = Bb + mu_j * dBdt;
  } else if (!has_surface) {
    Ibu = Bs;
  }

  // ---- the base below the last layer, then the base sweep ----
  Base<N> B;
  if (has_surface) {
    const double rr = 2.0 * A * mu_j * wt_j * a.xfac;
#pragma unroll
    for (int r = 0; r < RP; ++r) B.R[r] = (h * RP + r < a.n_real) ? rr : 0.0;
    double ssol = 0.0;
    if (a.has_solar && A > 0.0) ssol = (A / kPi) * a.F0 * a.mu0 * exp(-tau_tot / a.mu0);
    B.s = (1.0 - A) * Bs + ssol;
    B.scat = A > 0.0;
  } else {
#pragma unroll
    for (int r = 0; r < RP; ++r) B.R[r] = 0.0;
    B.s = Ibu;
    B.scat = 0;
  }
  for (int l = L - 1; l >= 0; --l)
    base_add<N>(B, lay + (long)l * REC, lflag[l], rbw + (long)l * F::size, j, h, sm);

  auto emit = [&](int l, double Iup, double Idn) {
    const double c = 2.0 * kPi * wt_j * mu_j;
    const double fu = group_sum_ordered<N, R>(c * Iup, j, h, sm + Lds<N>::pval);
    const double fdn = group_sum_ordered<N, R>(c * Idn, j, h, sm + Lds<N>::pval);
    double J = group_sum_ordered<N, R>(0.5 * wt_j * (Iup + Idn), j, h, sm + Lds<N>::pval);
    if (j == 0 && h == 0) {
      const long o = gw * (L + 1) + lev(a, l);
      double fdir = 0.0;
      if (a.has_solar) {
        fdir = (l == 0) ? a.F0 * a.mu0
                        : a.F0 * a.mu0 * exp(-a.tau_cum[(long)w * (L + 1) + l] / a.mu0);
        J += fdir * (1.0 / (4.0 * kPi * a.mu0));
      }
      a.fu[o] = fu * a.scaling;
      a.fd[o] = fdn;
      a.mi[o] = J;
      a.fdir[o] = fdir;
      if (a.fdiv) {
        const int lyr = (l == 0) ? 0 : l - 1;
        const double om = a.ssa[gw * L + lin(a, lyr)];
        a.fdiv[o] = 4.0 * kPi * (1.0 - om) * (J - Blev(l));
      }
    }
  };

  // ---- downward sweep (the interface intensities of src/solver.cpp:628-678) ----
  double Idn = Itd, Iup = B.s;
  if (B.scat) {
    put_col<N, R>(tile, j, h, B.R);
    if (h == 0) v0[j] = Itd;
    wave_sync();
    Iup += matvec<N, R>(tile, v0, j);
    wave_sync();
  }
  emit(0, Iup, Idn);
  for (int l = 0; l < L; ++l) {
    const double* o = rbw + (long)l * F::size;
    const double d = o[F::D + j], sbn = o[F::S + j];
    const bool bscat = o[F::F] != 0.0;
    if (h == 0) v0[j] = Idn;
    wave_sync();
    Idn = (lflag[l] ? gmatvec<N>(o + F::G, v0, j) : o[F::G + j] * Idn) + d;
    wave_sync();
    Iup = sbn;
    if (bscat) {
      if (h == 0) v0[j] = Idn;
      wave_sync();
      Iup = gmatvec<N>(o + F::R, v0, j) + sbn;
      wave_sync();
    }
    emit(l + 1, Iup, Idn);
  }
}

}  // namespace dev
}  // namespace adrt
