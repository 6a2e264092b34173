// adrt_adj16.hpp -- reverse pass (adjoint) of the N = 16 solve.
//
// The reference has no adjoint; this differentiates the same forward
// algorithm (src/solver.cpp:348-722, src/doubling.hpp:29-230,
// src/phase_matrix.hpp:61-181) in reverse mode:
//
//  1. adding16_kernel<true> solves the transposed interface system (see
//     adrt_fwd16.hpp) for the adjoint vectors: per interface l, D_l (the
//     multiplier of the equation of I_up(l)) and U_l (that of I_dn(l)).
//  2. layer16_adjoint_kernel, one wave per (wavenumber, layer): the
//     cotangents of the layer record are outer products of forward and adjoint
//     interface vectors,
//        R^ = D_l I_dn(l)^T + U_{l+1} I_up(l+1)^T,
//        T^ = D_l I_up(l+1)^T + U_{l+1} I_dn(l)^T,  s_up^ = D_l, s_dn^ = U_{l+1};
//     the doubling is recomputed from the thin-layer start with per-step
//     checkpoints of R, T and (I - R^2)^-1 in a per-wave slab, then walked
//     backwards with matrix-core products only; the thin-layer start, the
//     Hansen normalisation and the Legendre sums are then differentiated.
//  3. column16_adjoint_kernel, one thread per wavenumber: the cumulative
//     optical depth (solar beam at layer tops, direct flux, surface beam),
//     boundary and surface terms, level Planck values and the delta-M chain.
//  4. moments16_reduce_kernel sums moment gradients over wavenumbers when the
//     moments are shared.
#pragma once

#include "adrt_fwd16.hpp"

namespace adrt {
namespace dev {
namespace w16 {

struct AdjArgs {
  double* slab;   // per-wave checkpoints [nwaves][nn_cap][SLAB]
  int nwaves;     // waves of the persistent layer grid
  int nn_cap;     // doubling steps a slab holds
  int* status;    // bit 0: a layer needed more doubling steps than nn_cap
  int* queue;     // task counter of the persistent grid, zeroed before each launch
  // per-task results, chunk-local [W][L]
  double* g_taus;  // d/d scaled tau (layer terms; column terms added later)
  double* g_omgs;  // d/d scaled omega
  double* g_Bt;    // d/d B at the layer's top level
  double* g_Bb;    // d/d B at the layer's bottom level
  double* g_Ftop;  // d/d solar flux reaching the layer top
  double* g_f;     // d/d delta-M fraction through the scaled moments
  double* g_chi;   // [W][L][M] moment gradients (shared moments only)
  double* scot;    // [W][L][SCOT] cotangents of the doubling start (FP64 path)
  // gradients in the caller's layout; null = not wanted
  double* d_tau;     // [W][L]
  double* d_ssa;     // [W][L]
  double* d_mom;     // [W][L][M], or [L][M] for shared moments (accumulated over chunks)
  double* d_B;       // [W][L+1]
  double* d_surf;    // [W]
  double* d_top;     // [W]
  double* d_albedo;  // [W] per-wavenumber contributions
};

// The persistent grids take their tasks CHUNK at a time from a counter that
// the host zeroes before each launch: layers differ in doubling count (1 to
// ~20 steps), so a wave that drew short layers draws more instead of the grid
// waiting for the wave with the longest fixed share.  f(task) for every task
// below ntask, each once over the grid; the loop is wave-uniform and every
// wave leaves it once the counter has passed ntask.
template <int CHUNK, class F>
__device__ __forceinline__ void for_tasks(int* queue, long ntask, F f) {
  for (;;) {
    int t0 = 0;
    if ((threadIdx.x & 63) == 0) t0 = atomicAdd(queue, CHUNK);
    t0 = __builtin_amdgcn_readfirstlane(t0);
    if ((long)t0 >= ntask) break;
    const long t1 = (long)t0 + CHUNK < ntask ? (long)t0 + CHUNK : ntask;
    for (long task = t0; task < t1; ++task) f(task);
  }
}
// chunks: tasks of every layer; tasks of which most are skipped (the routed
// layers of the mixed form, the layers pair mode leaves)
constexpr int TASK_CHUNK = 4, SPARSE_TASK_CHUNK = 32;

// waves per SIMD of the persistent doubling-reverse grid (its register budget:
// 512 / 3 -> 168 registers per lane)
constexpr int REV_WAVES_PER_SIMD = 3;
constexpr int SLAB = 3 * N * N + 4 * N + 2;  // R, T, (I-R^2)^-1 (register order); y, z, s_up, s_dn; gamma
// cotangents of the doubling start: R0, T0 (register order); y0, s_up0, s_dn0;
// column sums of the g and gamma cotangents
constexpr int SCOT = 2 * N * N + 3 * N + 2;

struct LdsA {
  static constexpr int t0 = 0, t1 = TILE, t2 = 2 * TILE, t3 = 3 * TILE, vs = 4 * TILE;
  static constexpr int size = vs + 8 * N;  // 8 vector slots; the last one for column sums
};
static_assert(LdsA::t3 + TILE <= LdsA::vs && LdsA::vs + 8 * N <= LdsA::size,
              "reverse per-wave LDS regions overlap");
// start16_adjoint_kernel: per wave one tile, a column-sum slot and two tiles
constexpr int LDS_START16 = 3 * TILE + N;

// phase matrices and thin-layer scalars of a scattering layer, recomputed for
// the reverse pass in the register layout of RowT, keeping the sums before the
// Hansen normalisation (src/phase_matrix.hpp:61-181, src/doubling.hpp:83-147)
template <typename RowT>
struct Phase16 {
  double omega, con, f, inv1mf, csum, corr, qplus, qminus, ssol, corr_s, pplus, pminus;
  double tau0, Ftop, gam0;
  double Qpp[4], Qpm[4], Ppp[4], Ppm[4];  // rows row<RowT>(h, r) of column j
  const double* chi;
  int Lm, nn, mode;
  bool solar;
};

// Each wave's slab ends with a parking area for its lanes' Phase16: the phase
// data are needed again only after both doubling loops, so they wait in memory
// rather than in registers while the loops run.
constexpr int PHASE_WORDS = (int)((sizeof(Phase16<double>) + 7) / 8);
static_assert(sizeof(Phase16<double>) == sizeof(Phase16<float>), "one parking size");
__host__ __device__ constexpr long slab_stride(int nn_cap) {
  return (long)nn_cap * SLAB + 64 * PHASE_WORDS;
}
template <typename RowT>
__device__ __forceinline__ void park(double* slab, int nn_cap, const Lane& L, const Phase16<RowT>& p) {
  *(Phase16<RowT>*)(slab + (long)nn_cap * SLAB + L.l * PHASE_WORDS) = p;
}
template <typename RowT>
__device__ __forceinline__ void unpark(const double* slab, int nn_cap, const Lane& L, Phase16<RowT>& p) {
  p = *(const Phase16<RowT>*)(slab + (long)nn_cap * SLAB + L.l * PHASE_WORDS);
}

template <typename RowT>
__device__ __forceinline__ void phase16_recompute(const FwdArgs& a, long task, int w, int l,
                                                  long gw, const Lane& L, double tau,
                                                  double omega_in, int mode, double* dtile,
                                                  double* csl, Phase16<RowT>& p) {
  const int Lay = a.L, j = L.j;
  const double wt_j = a.wt[j];
  const int li = lin(a, l);
  p.mode = mode;
  p.omega = omega_in > 1.0 ? 1.0 : omega_in;
  p.con = 2.0 * p.omega * kPi;
  p.chi = a.shared_moments ? a.mom + (long)li * a.M : a.mom + (gw * Lay + li) * a.M;
  const int twoM = 2 * a.n_real;
  p.Lm = a.M;
  p.f = 0.0;
  p.inv1mf = 1.0;
  if (mode == 3) {
    p.f = a.fdm[task];
    p.inv1mf = 1.0 / (1.0 - p.f);
    p.Lm = twoM;
  } else if (mode == 2) {
    p.Lm = a.M < twoM ? a.M : twoM;
  }
  p.solar = a.has_solar != 0;
  if (a.phase) {
    // shared moments: the layer's phase record (phase16_kernel)
    const double* ph = a.phase + (long)l * PH16;
    p.csum = ph[PH_CSUM + j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      p.Qpp[r] = ph[row<RowT>(L.h, r) * N + j];
      p.Qpm[r] = ph[PH_QPM + row<RowT>(L.h, r) * N + j];
    }
    p.qplus = ph[PH_QP + j];
    p.qminus = ph[PH_QM + j];
    p.ssol = ph[PH_SSOL];
  } else {
    double spj, smj;
    phase16_sums<RowT>(a, p.chi, mode, p.f, L, p.Qpp, p.Qpm, spj, smj);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      p.Qpp[r] = p.Qpp[r] / (2.0 * kPi);
      p.Qpm[r] = p.Qpm[r] / (2.0 * kPi);
      dtile[j * TS + row<RowT>(L.h, r)] = (p.Qpp[r] + p.Qpm[r]) * a.wt[row<RowT>(L.h, r)];
    }
    wave_sync();
    p.csum = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) p.csum += dtile[j * TS + i];
    wave_sync();
    p.qplus = p.qminus = p.ssol = 0.0;
    if (p.solar) {
      p.qplus = spj / (2.0 * kPi);
      p.qminus = smj / (2.0 * kPi);
      p.ssol = colsum<double>((p.qplus + p.qminus) * wt_j, L, csl);
    }
  }
  p.corr = p.csum > 0.0 ? 1.0 / (2.0 * kPi * p.csum) : 1.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    p.Ppp[r] = p.csum > 0.0 ? p.Qpp[r] * p.corr : p.Qpp[r];
    p.Ppm[r] = p.csum > 0.0 ? p.Qpm[r] * p.corr : p.Qpm[r];
  }
  p.corr_s = 1.0;
  p.pplus = p.pminus = 0.0;
  if (p.solar) {
    p.corr_s = p.ssol > 0.0 ? 1.0 / (2.0 * kPi * p.ssol) : 1.0;
    p.pplus = p.ssol > 0.0 ? p.qplus * p.corr_s : p.qplus;
    p.pminus = p.ssol > 0.0 ? p.qminus * p.corr_s : p.qminus;
  }
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(p.omega);
  p.nn = nn0 < 1 ? 1 : nn0;
  p.tau0 = tau * ldexp(1.0, -p.nn);
  p.Ftop = 0.0;
  p.gam0 = 0.0;
  if (p.solar) {
    p.Ftop = a.F0 * exp(-a.tau_cum[(long)w * (Lay + 1) + l] / a.mu0);
    p.gam0 = exp(-p.tau0 / a.mu0);
  }
}

// The thin-layer start, the Hansen normalisation and the Legendre sums,
// reversed: from the cotangents of T0, R0 (rows row<RowT>(h, r)), y0 and the
// solar start vectors, and the column sums of the g and gamma cotangents, to
// the layer's tau, omega, solar-flux and moment gradients.
template <typename RowT>
__device__ __forceinline__ void start16_reverse(const FwdArgs& a, const Lane& L,
                                                const Phase16<RowT>& p, double omega_in,
                                                const double (&Rb0)[4], const double (&Tb0)[4],
                                                double yb, double susb, double sdsb,
                                                double g_sum, double gam_sum, double* csl,
                                                double* tiles, double* chi_out, double& g_tau,
                                                double& g_om, double& g_F, double& g_fr) {
  const int j = L.j;
  const double mu_j = a.mu[j], wt_j = a.wt[j];
  const double tau0 = p.tau0, con = p.con, omega = p.omega;
  double tb_l = 0.0, cb_l = 0.0, ob_l = 0.0, fb_l = 0.0;
  double Pb_pp[4], Pb_pm[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = row<RowT>(L.h, r);
    const double mui = a.mu[i];
    const double dij = (i == j) ? 1.0 : 0.0;
    const double tb = Tb0[r], rb = Rb0[r];
    tb_l += tb * (-(dij - con * p.Ppp[r] * wt_j) / mui) + rb * (con * p.Ppm[r] * wt_j / mui);
    cb_l += (tb * p.Ppp[r] + rb * p.Ppm[r]) * (tau0 * wt_j / mui);
    Pb_pp[r] = tb * (tau0 * con * wt_j / mui);
    Pb_pm[r] = rb * (tau0 * con * wt_j / mui);
  }
  double pplusb = 0.0, pminusb = 0.0;
  if (L.h == 0) {
    tb_l += yb * (1.0 - omega) / mu_j;
    ob_l += -yb * tau0 / mu_j;
  }
  if (p.solar) {
    const double base = omega * tau0 / mu_j * p.Ftop;
    const double baseb = susb * p.pminus + sdsb * p.pplus;
    pminusb = susb * base;
    pplusb = sdsb * base;
    if (L.h == 0) {
      ob_l += baseb * tau0 * p.Ftop / mu_j;
      tb_l += baseb * omega * p.Ftop / mu_j;
      fb_l += baseb * omega * tau0 / mu_j;
    }
  }
  double taub0 = wave_sum<double>(tb_l, L) + 0.5 * g_sum;
  if (p.solar) taub0 -= gam_sum * p.gam0 / a.mu0;
  const double conb = wave_sum<double>(cb_l, L);
  g_om = wave_sum<double>(ob_l, L) + 2.0 * kPi * conb;
  if (omega_in > 1.0) g_om = 0.0;  // clamped to 1
  g_F = p.solar ? wave_sum<double>(fb_l, L) : 0.0;
  g_tau += taub0 * ldexp(1.0, -p.nn);

  // Hansen normalisation and Legendre sums
  double Qb_pp[4], Qb_pm[4];
  if (p.csum > 0.0) {
    double q = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) q += Pb_pp[r] * p.Qpp[r] + Pb_pm[r] * p.Qpm[r];
    const double corrb = col_reduce<double>(q, L);
    const double csumb = -corrb * p.corr / p.csum;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Qb_pp[r] = Pb_pp[r] * p.corr + csumb * a.wt[row<RowT>(L.h, r)];
      Qb_pm[r] = Pb_pm[r] * p.corr + csumb * a.wt[row<RowT>(L.h, r)];
    }
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Qb_pp[r] = Pb_pp[r];
      Qb_pm[r] = Pb_pm[r];
    }
  }
  double qplusb = pplusb, qminusb = pminusb;
  if (p.solar && p.ssol > 0.0) {
    const double corrb = colsum<double>(pplusb * p.qplus + pminusb * p.qminus, L, csl);
    const double sb = -corrb * p.corr_s / p.ssol;
    qplusb = pplusb * p.corr_s + sb * wt_j;
    qminusb = pminusb * p.corr_s + sb * wt_j;
  }
  // Legendre sums on the matrix cores, 16 moments at a time: with
  // S+- = Q^++ +- Q^+- (rows in the FP64 instruction order) and the block
  // A(m, i) = P_{l0+m}(mu_i) in an LDS tile, U+- = A S+- holds
  // sum_i P_l(mu_i) S+-(i, j) for l = l0 + m, and
  //   chi^_l = (2l+1)/(2 pi) sum_j P_l(mu_j) (U(m, j) + P_l(mu0) w+-_j),
  // w+- = q+^ +- q-^, the sign + for even l; the sum over j is a DPP row sum,
  // so lane (j, h) ends with the moments l0 + h + 4r of its register r
  double* tP = tiles;
  double* tS = tiles + TILE;
  Mat<double> Sp, Sn;
  if constexpr (sizeof(RowT) == 8) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Sp.v[r] = Qb_pp[r] + Qb_pm[r];
      Sn.v[r] = Qb_pp[r] - Qb_pm[r];
    }
  } else {  // to the FP64 row order through LDS
#pragma unroll
    for (int r = 0; r < 4; ++r) tS[j * TS + row<RowT>(L.h, r)] = Qb_pp[r] + Qb_pm[r];
    wave_sync();
    get<double>(tS, L, Sp);
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) tS[j * TS + row<RowT>(L.h, r)] = Qb_pp[r] - Qb_pm[r];
    wave_sync();
    get<double>(tS, L, Sn);
    wave_sync();
  }
  const double wp = qplusb + qminusb, wn = qplusb - qminusb;
  const bool odd = (L.h & 1) != 0;  // parity of l0 + h + 4r
  double gfr_part = 0.0;
  for (int l0 = 0; l0 < p.Lm; l0 += 16) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ll = l0 + row<double>(L.h, r);
      tP[j * TS + row<double>(L.h, r)] = (ll < p.Lm) ? a.Pl[ll * N + j] : 0.0;
    }
    wave_sync();
    Mat<double> Up, Un;
    mul<double>(tP, Sp, L, Up);
    mul<double>(tP, Sn, L, Un);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = row<double>(L.h, r), ll = l0 + m;
      const bool live = ll < p.Lm;
      double u = odd ? Un.v[r] : Up.v[r];
      if (p.solar && live) u += a.Pl0[ll] * (odd ? wn : wp);
      const double tot = row_sum<double>(tP[j * TS + m] * u);
      if (live && j == 0) {
        const double chisb = tot * ((2 * ll + 1) / (2.0 * kPi));
        if (p.mode == 3) gfr_part += chisb * (p.chi[ll] - 1.0) * p.inv1mf * p.inv1mf;
        if (chi_out) chi_out[ll] = (p.mode == 3) ? chisb * p.inv1mf : chisb;
      }
    }
    wave_sync();
  }
  if (p.mode == 3) g_fr += wave_sum<double>(gfr_part, L);
}

// cotangents of a layer record from the interface vectors (see the file head)
struct LayerCot16 {
  const double* Dl;  // adjoint vector of the up-equation at l
  const double* U1;  // adjoint vector of the down-equation at l+1
  double idn_j, iup_j, sub, sdb;
};

__device__ __forceinline__ LayerCot16 layer_cot16(const FwdArgs& a, int w, int l, const Lane& L) {
  const int Lay = a.L, j = L.j;
  const double* Iw = a.Iv + (long)w * (Lay + 1) * 2 * N;
  const double* Aw = a.Av + (long)w * (Lay + 1) * 2 * N;
  LayerCot16 c;
  c.Dl = Aw + (long)l * 2 * N + N;
  c.U1 = Aw + (long)(l + 1) * 2 * N;
  c.idn_j = Iw[(long)l * 2 * N + N + j];
  c.iup_j = Iw[(long)(l + 1) * 2 * N + j];
  c.sub = c.Dl[j];
  c.sdb = c.U1[j];
  return c;
}

// R^ and T^ entries at row i of column j
__device__ __forceinline__ double cot_R(const LayerCot16& c, int i) {
  return c.Dl[i] * c.idn_j + c.U1[i] * c.iup_j;
}
__device__ __forceinline__ double cot_T(const LayerCot16& c, int i) {
  return c.Dl[i] * c.iup_j + c.U1[i] * c.idn_j;
}

// reverse of a transparent (no dependence) or purely absorbing layer
// (src/doubling.hpp:52-77); returns false for a scattering layer
__device__ __forceinline__ bool nonscat16_reverse(const FwdArgs& a, const Lane& L,
                                                  const LayerCot16& c, double tau,
                                                  double omega_in, double B_bar, double B_d,
                                                  double* csl, double& g_tau, double& g_Bt,
                                                  double& g_Bb) {
  if (tau <= 0.0) return true;
  if (omega_in > 0.0) return false;
  const int j = L.j;
  const double mu_j = a.mu[j];
  const double tex = -tau / mu_j;
  const bool live = tex > -200.0;
  const double t = live ? exp(tex) : 0.0;
  const double dt = live ? -t / mu_j : 0.0;
  const double omt = 1.0 - t;
  const double slope = mu_j * omt - 0.5 * tau * (1.0 + t);
  const double dslope = -mu_j * dt - 0.5 * (1.0 + t) - 0.5 * tau * dt;
  const double dBd = -B_d / tau;
  const double tbj = cot_T(c, j);
  const double dsu = -B_bar * dt + dBd * slope + B_d * dslope;
  const double dsd = -B_bar * dt - dBd * slope - B_d * dslope;
  g_tau = colsum<double>(tbj * dt + c.sub * dsu + c.sdb * dsd, L, csl);
  const double gbar = colsum<double>((c.sub + c.sdb) * omt, L, csl);
  const double gd = colsum<double>((c.sub - c.sdb) * slope, L, csl);
  g_Bt = 0.5 * gbar - gd / tau;
  g_Bb = 0.5 * gbar + gd / tau;
  return true;
}

__device__ void layer16_adjoint_task(const FwdArgs& a, const AdjArgs& b, long task, const Lane& L,
                                     double* sm, double* slab) {
  using R = double;
  const int Lay = a.L;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  const int j = L.j;
  R* t0 = sm + LdsA::t0;
  R* t1 = sm + LdsA::t1;
  R* t2 = sm + LdsA::t2;
  R* t3 = sm + LdsA::t3;
  R* vs = sm + LdsA::vs;
  R* csl = vs + 7 * N;  // column-sum slot

  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  const int mode = a.mode[task];
  const double Btop = a.B ? a.B[gw * (Lay + 1) + lev(a, l)] : 0.0;
  const double Bbot = a.B ? a.B[gw * (Lay + 1) + lev(a, l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (tau > 0.0) ? (Bbot - Btop) / tau : 0.0;
  const int li = lin(a, l);
  double* chi_out = a.shared_moments ? b.g_chi + task * a.M
                                     : (b.d_mom ? b.d_mom + (gw * Lay + li) * a.M : nullptr);
  const LayerCot16 c = layer_cot16(a, w, l, L);
  double g_tau = 0.0, g_om = 0.0, g_Bt = 0.0, g_Bb = 0.0, g_F = 0.0, g_fr = 0.0;
  auto finish = [&](int nchi) {  // moments nchi..M-1 get zero
    if (chi_out)
      for (int m = nchi + L.l; m < a.M; m += 64) chi_out[m] = 0.0;
    if (L.l == 0) {
      b.g_taus[task] = g_tau;
      b.g_omgs[task] = g_om;
      b.g_Bt[task] = g_Bt;
      b.g_Bb[task] = g_Bb;
      b.g_Ftop[task] = g_F;
      b.g_f[task] = g_fr;
    }
  };
  if (nonscat16_reverse(a, L, c, tau, omega_in, B_bar, B_d, csl, g_tau, g_Bt, g_Bb)) {
    finish(0);
    return;
  }
  Mat<R> Rb, Tb;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Rb.v[r] = cot_R(c, row<R>(L.h, r));
    Tb.v[r] = cot_T(c, row<R>(L.h, r));
  }
  const double sub = c.sub, sdb = c.sdb;

  // the thin-layer start as the forward computed it; its reverse (phase
  // matrices, Hansen normalisation, Legendre sums) is start16_adjoint_kernel's
  Mat<R> Rm, Tm;
  double y, z = 0.0, sus, sds, gam, gk, tau0;
  int nn;
  {
    Start16<R> st0;
    layer16_start<R>(a, task, w, l, gw, L, tau, omega_in, mode, t0, csl, st0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Tm.v[r] = st0.Tv[r];
      Rm.v[r] = st0.Rv[r];
    }
    y = st0.y;
    sus = st0.sus;
    sds = st0.sds;
    gam = st0.gam;
    gk = st0.gk;
    tau0 = st0.tau0;
    nn = st0.nn;
  }
  if (nn > b.nn_cap) {
    if (L.l == 0) atomicOr(b.status, 1);
    finish(0);
    return;
  }
  const bool solar = a.has_solar != 0;
  int rows[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) rows[r] = row<R>(L.h, r);
  R* s0 = vs;
  R* s1 = vs + N;
  R* s2 = vs + 2 * N;
  R* s3 = vs + 3 * N;
  R* s4 = vs + 4 * N;
  R* s5 = vs + 5 * N;
  R* s6 = vs + 6 * N;

  // ---- doubling recomputed with checkpoints; (I-R^2)^-1 by Newton-Schulz ----
  for (int k = 0; k < nn; ++k) {
    R* st = slab + (long)k * SLAB;
    gstore<R>(st, L, Rm);
    gstore<R>(st + N * N, L, Tm);
    if (L.h == 0) {
      st[3 * N * N + j] = y;
      st[3 * N * N + N + j] = z;
      st[3 * N * N + 2 * N + j] = sus;
      st[3 * N * N + 3 * N + j] = sds;
    }
    if (L.l == 0) st[3 * N * N + 4 * N] = gam;
    Mat<R> A, Ai, TG, Y, X;
    put<R>(t0, L, Rm);
    if (L.h == 0) {
      s0[j] = sds;
      s1[j] = sus;
    }
    wave_sync();
    mul<R>(t0, Rm, L, A);  // P = R^2
    double b1 = 0.0, b2 = 0.0;
    if (solar) {
      double bs[2];
      matvecs<R, 2>(t0, s0, L, bs);
      b1 = bs[0];
      b2 = bs[1];
    }
    wave_sync();
    put<R>(t0, L, A);
    wave_sync();
    inverse_ns<R>(A, t0, t3, Ai, L);  // (I - R^2)^-1
    gstore<R>(st + 2 * N * N, L, Ai);
    put<R>(t1, L, Tm);
    wave_sync();
    mul<R>(t1, Ai, L, TG);  // TG = T (I - R^2)^-1
    put<R>(t2, L, TG);
    if (L.h == 0) {
      s0[j] = z + gk * y;
      s1[j] = y;
      s2[j] = b1 + gam * sus;
      s3[j] = gam * b2 + sds;
    }
    wave_sync();
    const bool last = k == nn - 1;  // the last R_new, T_new are never read back
    mul<R>(t2, Rm, L, Y);  // TGR
    if (!last) mul<R>(t2, Tm, L, X);  // T_new = TG T
    double a1, a3, c1 = 0.0, c2 = 0.0;
    if (solar) {
      double s4v[4];
      matvecs<R, 4>(t2, s0, L, s4v);
      a1 = s4v[0];
      a3 = s4v[1];
      c1 = s4v[2];
      c2 = s4v[3];
    } else {
      double s2v[2];
      matvecs<R, 2>(t2, s0, L, s2v);
      a1 = s2v[0];
      a3 = s2v[1];
    }
    put<R>(t0, L, Y);
    wave_sync();
    if (!last) mul_add<R>(t0, Tm, Rm, L, Rm);  // R_new = R + TGR T
    double s2v[2];
    matvecs<R, 2>(t0, s0, L, s2v);
    wave_sync();
    if (!last) Tm = X;
    const double znew = (a1 - s2v[0]) + z - gk * y;
    const double ynew = a3 + s2v[1] + y;
    z = znew;
    y = ynew;
    if (solar) {
      sus = c1 + sus;
      sds = c2 + gam * sds;
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // ---- layer assembly, reversed: s_up = y B_bar + z B_d + s_up_sol, s_dn = y B_bar - z B_d + s_dn_sol
  double yb = (sub + sdb) * B_bar, zb = (sub - sdb) * B_d, susb = sub, sdsb = sdb;
  {
    const double gbar = colsum<R>((sub + sdb) * y, L, csl);
    const double gd = colsum<R>((sub - sdb) * z, L, csl);
    g_Bt = 0.5 * gbar - gd / tau;
    g_Bb = 0.5 * gbar + gd / tau;
    g_tau = -gd * B_d / tau;
  }
  double gamb = 0.0, gb = 0.0;  // per-column parts of the scalar cotangents
  Mat<R> Rbar = Rb, Tbar = Tb;

  // ---- doubling steps in reverse ----
  for (int k = nn - 1; k >= 0; --k) {
    const R* st = slab + (long)k * SLAB;
    Mat<R> Rk, Tk, Ai, TG, TGR, Tt, Rt;
    gload<R>(st, L, Rk);
    gload<R>(st + N * N, L, Tk);
    gload<R>(st + 2 * N * N, L, Ai);
    const double yk = st[3 * N * N + j], zk = st[3 * N * N + N + j];
    const double susk = st[3 * N * N + 2 * N + j], sdsk = st[3 * N * N + 3 * N + j];
    const double gm = st[3 * N * N + 4 * N];
    const double g = 0.5 * tau0 * ldexp(1.0, k);
    put<R>(t0, L, Tk);
    put<R>(t2, L, Rk);
    if (L.h == 0) {
      s0[j] = zb;
      s1[j] = yb;
      s2[j] = susb;
      s3[j] = sdsb;
      s4[j] = sdsk;
      s5[j] = susk;
    }
    wave_sync();
    mul<R>(t0, Ai, L, TG);
    get_t<R>(t0, L, Tt);
    get_t<R>(t2, L, Rt);
    wave_sync();
    put<R>(t1, L, TG);
    wave_sync();
    mul<R>(t1, Rk, L, TGR);
    const double u = zk + g * yk;
    // vector recursions, reversed
    const double x_zb = tmatvec<R>(TG, s0, L);
    const double x_yb = tmatvec<R>(TG, s1, L);
    const double r_zb = tmatvec<R>(TGR, s0, L);
    const double r_yb = tmatvec<R>(TGR, s1, L);
    const double ub = x_zb - r_zb;
    const double yb_n = ((yb - g * zb) + (x_yb + r_yb)) + g * ub;
    const double zb_n = zb + ub;
    gb = 2.0 * gb + (ub - zb) * yk;
    double v2 = 0.0, v3 = 0.0, b1b = 0.0, b2b = 0.0, susb_n = susb, sdsb_n = sdsb;
    if (solar) {
      const double v2b = tmatvec<R>(TG, s2, L);
      const double v3b = tmatvec<R>(TG, s3, L);
      double bs[2];
      matvecs<R, 2>(t2, s4, L, bs);  // R s_dn, R s_up
      v2 = bs[0] + gm * susk;
      v3 = gm * bs[1] + sdsk;
      b1b = v2b;
      b2b = gm * v3b;
      wave_sync();
      if (L.h == 0) {
        s6[j] = b1b;
        s0[j] = b2b;  // s0 is free again (read above)
      }
      wave_sync();
      const double rt_b1 = tmatvec<R>(Rk, s6, L);
      const double rt_b2 = tmatvec<R>(Rk, s0, L);
      susb_n = (susb + gm * v2b) + rt_b2;
      sdsb_n = (gm * sdsb + v3b) + rt_b1;
      gamb = 2.0 * gm * gamb + ((sdsb * sdsk + v2b * susk) + v3b * bs[1]);
    }
    // matrices, reversed
    vec4<R> acc = zero4<R>();
    put<R>(t0, L, Rbar);
    put<R>(t3, L, Tbar);
    wave_sync();
    outer4<R>(acc, -zb, yb, 0.0, 0.0, u, yk, 0.0, 0.0, L);
    mac<R>(acc, t0, Tt, L);  // TGR^ = R'^ T^T + vector terms
    const Mat<R> TGRb = mat_of<R>(acc);
    acc = zero4<R>();
    outer4<R>(acc, zb, yb, susb, sdsb, u, yk, v2, v3, L);
    mac<R>(acc, t3, Tt, L);  // TG^ = T'^ T^T + ...
    wave_sync();
    put<R>(t0, L, TGRb);
    wave_sync();
    mac<R>(acc, t0, Rt, L);  // ... + TGR^ R^T
    const Mat<R> TGb = mat_of<R>(acc);
    wave_sync();
    put<R>(t3, L, Ai);
    wave_sync();
    Mat<R> AiT;
    get_t<R>(t3, L, AiT);
    wave_sync();
    put<R>(t3, L, TGb);
    wave_sync();
    acc = zero4<R>();
    mac<R>(acc, t3, AiT, L);  // W = TG^ (I - R^2)^-T
    const Mat<R> W = mat_of<R>(acc);
    acc = zero4<R>();
    mac_t<R>(acc, t1, W, L);  // Abn = -A^ = TG^T W
    const Mat<R> Abn = mat_of<R>(acc);
    acc = acc_of<R>(Rbar);
    outer4<R>(acc, b1b, b2b, 0.0, 0.0, sdsk, susk, 0.0, 0.0, L);
    mac_t<R>(acc, t1, TGRb, L);  // + TG^T TGR^
    wave_sync();
    put<R>(t0, L, Abn);
    wave_sync();
    mac<R>(acc, t0, Rt, L);     // + Abn R^T
    mac_t<R>(acc, t2, Abn, L);  // + R^T Abn
    const Mat<R> Rbar_n = mat_of<R>(acc);
    wave_sync();
    put<R>(t3, L, TGR);
    wave_sync();
    acc = acc_of<R>(W);
    mac_t<R>(acc, t3, Rbar, L);  // + TGR^T R'^
    mac_t<R>(acc, t1, Tbar, L);  // + TG^T T'^
    Tbar = mat_of<R>(acc);
    Rbar = Rbar_n;
    yb = yb_n;
    zb = zb_n;
    susb = susb_n;
    sdsb = sdsb_n;
    wave_sync();
  }

  // ---- cotangents of the doubling start, for start16_adjoint_kernel ----
  const double g_sum = colsum<R>(gb, L, csl);
  const double gam_sum = solar ? colsum<R>(gamb, L, csl) : 0.0;
  double* sc = b.scot + task * SCOT;
  gstore<R>(sc, L, Rbar);
  gstore<R>(sc + N * N, L, Tbar);
  if (L.h == 0) {
    sc[2 * N * N + j] = yb;
    sc[2 * N * N + N + j] = susb;
    sc[2 * N * N + 2 * N + j] = sdsb;
  }
  if (L.l == 0) {
    sc[2 * N * N + 3 * N] = g_sum;
    sc[2 * N * N + 3 * N + 1] = gam_sum;
    b.g_taus[task] = g_tau;  // completed by start16_adjoint_kernel
    b.g_Bt[task] = g_Bt;
    b.g_Bb[task] = g_Bb;
  }
}

// The thin-layer start, the Hansen normalisation and the Legendre sums of
// each scattering layer, reversed, from the start cotangents that
// layer16_adjoint_kernel left in b.scot; one wave per (wavenumber, layer).
// Kept apart from the doubling reverse so that kernel keeps to 128 registers.
// 4 waves per SIMD (measured: 24.0 ms on the 65536 x 100 workload against
// 27.6 ms at 3 and 32.4 ms at 5).
// RowT = float: a layer of the mixed form's FP32 reverse (start cotangents in
// the FP32 register order of layer16m_adjoint_kernel).
template <typename RowT>
__device__ __forceinline__ void start16_adjoint_task(const FwdArgs& a, const AdjArgs& b, long task,
                                                     const Lane& L, double* dtile) {
  const int Lay = a.L;
  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  if (tau <= 0.0 || omega_in <= 0.0) return;  // finished by the doubling reverse
  const int j = L.j;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  double* csl = dtile + TILE;
  double* tiles = csl + N;  // two tiles for the Legendre sums
  Phase16<RowT> p;
  phase16_recompute<RowT>(a, task, w, l, gw, L, tau, omega_in, a.mode[task], dtile, csl, p);
  if (p.nn > b.nn_cap) return;  // reported through b.status
  const double* sc = b.scot + task * SCOT;
  double Rb0[4], Tb0[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Rb0[r] = sc[L.l * 4 + r];
    Tb0[r] = sc[N * N + L.l * 4 + r];
  }
  const double yb = sc[2 * N * N + j], susb = sc[2 * N * N + N + j];
  const double sdsb = sc[2 * N * N + 2 * N + j];
  const double g_sum = sc[2 * N * N + 3 * N], gam_sum = sc[2 * N * N + 3 * N + 1];
  const int li = lin(a, l);
  double* chi_out = a.shared_moments ? b.g_chi + task * a.M
                                     : (b.d_mom ? b.d_mom + (gw * Lay + li) * a.M : nullptr);
  double g_tau = b.g_taus[task], g_om = 0.0, g_F = 0.0, g_fr = 0.0;
  start16_reverse<RowT>(a, L, p, omega_in, Rb0, Tb0, yb, susb, sdsb, g_sum, gam_sum, csl,
                        tiles, chi_out, g_tau, g_om, g_F, g_fr);
  if (chi_out)
    for (int m = p.Lm + L.l; m < a.M; m += 64) chi_out[m] = 0.0;
  if (L.l == 0) {
    b.g_taus[task] = g_tau;
    b.g_omgs[task] = g_om;
    b.g_Ftop[task] = g_F;
    b.g_f[task] = g_fr;
  }
}

// MIXED: the layers of the mixed_reverse form, in one launch: those the
// forward routed to the FP64 kernels (redo) in the FP64 register order of
// layer16_adjoint_kernel<true>, the others in the FP32 order.
template <bool MIXED = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
start16_adjoint_kernel(FwdArgs a, AdjArgs b) {
  __shared__ double smem[4 * LDS_START16];
  const int wv = wave_index();
  const long task = (long)blockIdx.x * 4 + wv;
  if (task >= (long)a.W * a.L) return;
  const Lane L;
  double* dtile = smem + wv * LDS_START16;
  if constexpr (MIXED) {
    if (a.redo[task])
      start16_adjoint_task<double>(a, b, task, L, dtile);
    else
      start16_adjoint_task<float>(a, b, task, L, dtile);
  } else {
    start16_adjoint_task<double>(a, b, task, L, dtile);
  }
}

// REDO: only the layers the mixed-precision forward left to the FP64 kernel;
// UNPAIRED: only the layers pair mode leaves (a.pflag, adrt_pair16.hpp)
template <bool REDO, bool UNPAIRED = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(REV_WAVES_PER_SIMD, 8)))
layer16_adjoint_kernel(FwdArgs a, AdjArgs b) {
  __shared__ double smem[4 * LdsA::size];
  const int wv = wave_index();
  const int gwave = blockIdx.x * 4 + wv;
  const Lane L;
  double* sm = smem + wv * LdsA::size;
  double* slab = b.slab + (long)gwave * slab_stride(b.nn_cap);
  const long ntask = (long)a.W * a.L;
  for_tasks<(REDO || UNPAIRED) ? SPARSE_TASK_CHUNK : TASK_CHUNK>(b.queue, ntask, [&](long task) {
    if constexpr (REDO) {
      if (!a.redo[task]) return;
    }
    if constexpr (UNPAIRED) {
      if (a.pflag[task]) return;
    }
    layer16_adjoint_task(a, b, task, L, sm, slab);
  });
}

// the column terms of the reverse pass for interface vectors of NV elements;
// one thread per wavenumber
template <int NV>
__device__ __forceinline__ void column_adjoint(const FwdArgs& a, const AdjArgs& b, int w) {
  const long gw = (long)(a.w0 + w);
  const int Lay = a.L;
  const double* Aw = a.Av + (long)w * (Lay + 1) * 2 * NV;
  const double* Iw = a.Iv + (long)w * (Lay + 1) * 2 * NV;
  double* gts = b.g_taus + (long)w * Lay;
  const double* Bw = a.B ? a.B + gw * (Lay + 1) : nullptr;
  auto Blev = [&](int l) -> double { return Bw ? Bw[lev(a, l)] : 0.0; };
  const double Bs = a.surf_w ? a.surf_w[gw] : a.surf_em;
  const double A = a.albedo;
  const int has_surface = !a.diffusion && (A > 0.0 || Bs > 0.0);
  const double tau_tot = a.tau_cum[(long)w * (Lay + 1) + Lay];
  const double* DL = Aw + (long)Lay * 2 * NV + NV;
  double sumD = 0.0, sumDn = 0.0, sumU0 = 0.0, muwI = 0.0, muD = 0.0;
  for (int i = 0; i < NV; ++i) {
    sumD += DL[i];
    if (i < a.n_real) sumDn += DL[i];
    sumU0 += Aw[i];
    muwI += a.mu[i] * a.wt[i] * Iw[(long)Lay * 2 * NV + NV + i];
    muD += DL[i] * a.mu[i];
  }
  // cumulative optical depth: solar beam at layer tops, direct flux and beam
  // term of J at interfaces, surface beam; tau_cum(l) = sum of scaled tau above l
  if (a.has_solar) {
    double run = 0.0;
    for (int l = Lay; l >= 1; --l) {
      const double ex = exp(-a.tau_cum[(long)w * (Lay + 1) + l] / a.mu0);
      const long o = gw * (Lay + 1) + lev(a, l);
      double cdir = a.cdir ? a.cdir[o] : 0.0;
      double cJ = a.cJ ? a.cJ[o] : 0.0;
      if (a.cdiv) cJ += a.cdiv[o] * (4.0 * kPi * (1.0 - a.ssa[gw * Lay + lin(a, l - 1)]));
      cdir += cJ / (4.0 * kPi * a.mu0);
      double tcb = -cdir * a.F0 * ex;                                   // F_dir = F0 mu0 ex
      if (l < Lay) tcb -= b.g_Ftop[(long)w * Lay + l] * a.F0 * ex / a.mu0;  // F_top = F0 ex
      if (l == Lay && has_surface && A > 0.0) tcb -= sumD * (A / kPi) * a.F0 * ex;
      run += tcb;
      gts[l - 1] += run;
    }
  }
  // boundaries (src/solver.cpp:385-419, 516-551)
  if (b.d_top) b.d_top[gw] = sumU0;
  double dBL = 0.0, dBLm1 = 0.0;
  if (a.diffusion) {
    const double dt = a.tau_s[(long)w * Lay + Lay - 1];
    if (dt > 0.0) {
      dBL = sumD + muD / dt;
      dBLm1 = -muD / dt;
      gts[Lay - 1] += -muD * (Blev(Lay) - Blev(Lay - 1)) / (dt * dt);
    } else {
      dBL = sumD;
    }
    if (b.d_surf) b.d_surf[gw] = 0.0;
    if (b.d_albedo) b.d_albedo[gw] = 0.0;
  } else {
    if (b.d_surf) b.d_surf[gw] = (1.0 - A) * sumD;
    if (b.d_albedo) {
      double g = 2.0 * a.xfac * sumDn * muwI - Bs * sumD;
      if (a.has_solar) g += sumD * a.F0 * a.mu0 * exp(-tau_tot / a.mu0) / kPi;
      b.d_albedo[gw] = g;
    }
  }
  // level Planck values
  if (b.d_B) {
    for (int l = 0; l <= Lay; ++l) {
      double v = 0.0;
      if (l >= 1) v += b.g_Bb[(long)w * Lay + l - 1];
      if (l < Lay) v += b.g_Bt[(long)w * Lay + l];
      const long o = gw * (Lay + 1) + lev(a, l);
      if (a.cdiv) {
        const int lyr = (l == 0) ? 0 : l - 1;
        v -= a.cdiv[o] * (4.0 * kPi * (1.0 - a.ssa[gw * Lay + lin(a, lyr)]));
      }
      if (l == Lay) v += dBL;
      if (l == Lay - 1) v += dBLm1;
      b.d_B[o] = v;
    }
  }
  // divergence omega, delta-M chain (src/solver.cpp:445-486)
  for (int l = 0; l < Lay; ++l) {
    const long t = (long)w * Lay + l;
    const int li = lin(a, l);
    const double taus_b = gts[l], oms_b = b.g_omgs[t];
    const double tau = a.tau[gw * Lay + li], om = a.ssa[gw * Lay + li];
    double omb = 0.0, taub;
    if (a.cdiv) {
      const long o1 = gw * (Lay + 1) + lev(a, l + 1);
      omb -= a.cdiv[o1] * 4.0 * kPi * (a.mi[o1] - Blev(l + 1));
      if (l == 0) {
        const long o0 = gw * (Lay + 1) + lev(a, 0);
        omb -= a.cdiv[o0] * 4.0 * kPi * (a.mi[o0] - Blev(0));
      }
    }
    if (a.mode[t] == 3) {
      const double f = a.fdm[t], d = 1.0 - om * f;
      taub = taus_b * d;
      omb += oms_b * (1.0 - f) / (d * d) - taus_b * f * tau;
      const double fb = b.g_f[t] + oms_b * om * (om - 1.0) / (d * d) - taus_b * om * tau;
      double* co = a.shared_moments ? b.g_chi + t * a.M
                                    : (b.d_mom ? b.d_mom + (gw * Lay + li) * a.M : nullptr);
      if (co) co[2 * a.n_real] += fb;
    } else {
      taub = taus_b;
      omb += oms_b;
    }
    if (b.d_tau) b.d_tau[gw * Lay + li] = taub;
    if (b.d_ssa) b.d_ssa[gw * Lay + li] = omb;
  }
}

__global__ void __launch_bounds__(256) column16_adjoint_kernel(FwdArgs a, AdjArgs b) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < a.W) column_adjoint<N>(a, b, w);
}

// shared moments: d_mom[l'][m] += sum over the chunk's wavenumbers, in a fixed
// order (deterministic): MOM_SLICES partial sums per (layer, moment) over
// contiguous wavenumber slices (coalesced over m), then the slices in order.
constexpr int MOM_SLICES = 64;

__global__ void __launch_bounds__(256) moments16_partial_kernel(FwdArgs a, AdjArgs b,
                                                                double* part) {
  const int l = blockIdx.x, sl = blockIdx.y;
  const int tm = threadIdx.x & 63, tw = threadIdx.x >> 6;
  const int per = (a.W + MOM_SLICES - 1) / MOM_SLICES;
  const int wb = sl * per, we = min(a.W, wb + per);
  __shared__ double red[4][64];
  for (int m0 = 0; m0 < a.M; m0 += 64) {
    const int m = m0 + tm;
    double s = 0.0;
    if (m < a.M)
      for (int w = wb + tw; w < we; w += 4) s += b.g_chi[((long)w * a.L + l) * a.M + m];
    red[tw][tm] = s;
    __syncthreads();
    if (tw == 0 && m < a.M)
      part[((long)l * MOM_SLICES + sl) * a.M + m] = ((red[0][tm] + red[1][tm]) + red[2][tm]) + red[3][tm];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) moments16_reduce_kernel(FwdArgs a, AdjArgs b,
                                                               const double* part) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)a.L * a.M) return;
  const int l = (int)(k / a.M), m = (int)(k % a.M);
  double s = 0.0;
  for (int sl = 0; sl < MOM_SLICES; ++sl) s += part[((long)l * MOM_SLICES + sl) * a.M + m];
  b.d_mom[(long)lin(a, l) * a.M + m] += s;
}

__global__ void fill_kernel(double* p, long n, double v) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

__global__ void fill_int_kernel(int* p, int v) { *p = v; }

}  // namespace w16
}  // namespace dev
}  // namespace adrt
