// adrt_pair16.hpp -- two layers of quadrature order <= 8 per wave (gfx950).
//
// The reverse pass runs on the order-16 kernels, lower orders padded with
// zero-weight nodes (mu = 1, w = 0, P_l = 0).  A padded layer's matrices are
// block diagonal -- the real nodes 0..n-1 never couple to the pad nodes -- so
// only the real block matters.  Pair mode puts the real blocks of two layers
// into one 16 x 16 block-diagonal problem: layer A in
// rows and columns 0..7, layer B in 8..15.  Products then take two matrix
// instructions instead of four (pmul, adrt_wave16.hpp), and one wave does the
// doubling of two layers, a quarter of the matrix-core work per layer of the
// padded form.  Each lane's half (column j < 8 or >= 8) selects the layer
// whose scalars (tau0, its doubling count, the beam attenuation, ...) it
// carries; the doubling runs for the larger of the two counts, and a layer
// whose count is reached is stored and frozen (its matrices zeroed, its
// cotangents passed through unchanged in the reverse walk).
//
// Records leave in the padded order-16 layout: the real block in rows and
// columns 0..7, zeros elsewhere (pad nodes then carry no intensity at all;
// their weights are zero, so no output sees them).  The pairs come from
// pair_order_kernel (a.pord): each column's layers sorted by doubling count and
// taken two by two, so the layers of a pair double nearly as often.  A pair
// with a layer that does not scatter, and the last layer of an odd column, run
// one layer per wave in the single-layer kernels, launched after the pair
// kernels with UNPAIRED set (a.pflag = 0).
#pragma once

#include "adrt_adj16.hpp"

namespace adrt {
namespace dev {
namespace w16 {

// the lane's layer of the pair and the node of its column within that layer
struct PairLane {
  int hb;     // 0: layer A (rows and columns 0..7), 1: layer B (8..15)
  int jj;     // j & 7
  long task;  // task of this lane's layer
  int w, l;
  long gw;
  __device__ PairLane(const FwdArgs& a, const Lane& L, long tA, long tB) {
    hb = pair_half(L);
    jj = L.j & 7;
    task = hb ? tB : tA;
    w = (int)(task / a.L);
    l = (int)(task % a.L);
    gw = (long)(a.w0 + w);
  }
};

// sum over the 8 columns of the lane's layer, in column order (the order of
// colsum over a padded column, whose pad entries add exact zeros)
__device__ __forceinline__ double pair_colsum(double v, const Lane& L, double* slot) {
  if (L.h == 0) slot[L.j] = v;
  wave_sync();
  const int base = pair_half(L) * 8;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += slot[base + i];
  wave_sync();
  return s;
}

// a wave-uniform copy of the lanes' value for layer A (lane 0) and layer B (lane 8)
__device__ __forceinline__ int pair_max(int v) {
  const int va = __builtin_amdgcn_readlane(v, 0), vb = __builtin_amdgcn_readlane(v, 8);
  return va > vb ? va : vb;
}

// Pair order, one wave per column: each layer's stable rank by doubling count
// (non-scattering layers first, key -1), pord[rank] = layer, and the flag of
// the pair (ranks 2q, 2q + 1) it belongs to: it runs iff both layers scatter;
// the last rank of an odd column runs alone.  Columns of up to PAIR_MAX_L
// layers (the host enables pair mode only for those).
constexpr int PAIR_MAX_L = 1024;
__global__ void __launch_bounds__(256) pair_order_kernel(FwdArgs a) {
  __shared__ int lds[4][2][PAIR_MAX_L];
  const int wv = wave_index();
  const long w = (long)blockIdx.x * 4 + wv;
  if (w >= a.W) return;
  const int L = a.L, lane = threadIdx.x & 63;
  int* key = lds[wv][0];   // key of layer l
  int* rkey = lds[wv][1];  // key at rank r
  for (int l = lane; l < L; l += 64) {
    const long t = w * L + l;
    key[l] = a.mode[t] != 0 ? doubling_count(a.tau_s[t], a.omega_s[t]) : -1;
  }
  wave_sync();
  for (int l = lane; l < L; l += 64) {
    const int k = key[l];
    int r = 0;
    for (int m = 0; m < L; ++m) {
      const int km = key[m];
      r += (km < k || (km == k && m < l)) ? 1 : 0;
    }
    a.pord[w * L + r] = l;
    rkey[r] = k;
  }
  wave_sync();
  for (int l = lane; l < L; l += 64) {
    const int k = key[l];
    int r = 0;
    for (int m = 0; m < L; ++m) {
      const int km = key[m];
      r += (km < k || (km == k && m < l)) ? 1 : 0;
    }
    const int rp = r ^ 1;
    a.pflag[w * L + l] = (rp < L && k >= 0 && rkey[rp] >= 0) ? 1 : 0;
  }
}

// the tasks of pair p of a chunk (pair q = p mod L/2 of column w = p div L/2:
// layers pord[2q], pord[2q + 1]); false if the pair does not run (a layer of
// it does not scatter)
__device__ __forceinline__ bool pair_tasks(const FwdArgs& a, long p, long& tA, long& tB) {
  const int half = a.L / 2;
  const long w = p / half;
  const int q = (int)(p % half);
  const int* ord = a.pord + w * a.L;
  tA = w * a.L + ord[2 * q];
  tB = w * a.L + ord[2 * q + 1];
  return a.pflag[tA] != 0;
}

// Stores this lane's part of a block-diagonal pair matrix as its layer's
// padded order-16 matrix in register order: the real entry (ii, jj) at record
// lane jj + 16 h, register ii >> 2 (ii = row(h, r) - 8 hb), zeros for the pad
// rows of column jj and for the pad column jj + 8.
__device__ __forceinline__ void pair_gstore(double* g, const Lane& L, const PairLane& P,
                                            const Mat<double>& m) {
  double* c = g + (P.jj + 16 * L.h) * 4;
  double* z = g + (P.jj + 8 + 16 * L.h) * 4;
  c[0] = m.v[2 * P.hb];
  c[1] = m.v[2 * P.hb + 1];
  c[2] = 0.0;
  c[3] = 0.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) z[r] = 0.0;
}
// element jj of a vector of the lane's layer (lanes h = 0), zeros for the pad
// elements (lanes h = 1)
__device__ __forceinline__ void pair_vstore(double* g, const Lane& L, const PairLane& P, double v) {
  if (L.h == 0) g[P.jj] = v;
  if (L.h == 1) g[P.jj + 8] = 0.0;
}

// phase matrices and thin-layer start of the lane's layer, in the pair
// layout (layer16_start for two layers: src/phase_matrix.hpp:61-181,
// src/doubling.hpp:83-147); entries outside the lane's block are zero
__device__ __forceinline__ void pair_start(const FwdArgs& a, const Lane& L, const PairLane& P,
                                           double tau, double omega_in, int mode, double* dtile,
                                           double* pv, Start16<double>& st) {
  const int L_ = a.L, j = L.j, jj = P.jj;
  const double mu_j = a.mu[jj], wt_j = a.wt[jj];
  const double omega = omega_in > 1.0 ? 1.0 : omega_in;
  const double con = 2.0 * omega * kPi;
  const int li = lin(a, P.l);
  int rows[4], ii[4];
  bool inb[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    rows[r] = row<double>(L.h, r);
    ii[r] = rows[r] & 7;
    inb[r] = (rows[r] >> 3) == P.hb;
  }
  double Ppp[4], Ppm[4];
  double pplus = 0.0, pminus = 0.0;
  if (a.phase) {
    const double* ph = a.phase + (long)P.l * PH16;
    const double csum = ph[PH_CSUM + jj];
    const double corr = csum > 0.0 ? 1.0 / (2.0 * kPi * csum) : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Ppp[r] = inb[r] ? ph[ii[r] * N + jj] * corr : 0.0;
      Ppm[r] = inb[r] ? ph[PH_QPM + ii[r] * N + jj] * corr : 0.0;
    }
    if (a.has_solar) {
      pplus = ph[PH_QP + jj];
      pminus = ph[PH_QM + jj];
      const double s = ph[PH_SSOL];
      if (s > 0.0) {
        const double c = 1.0 / (2.0 * kPi * s);
        pplus *= c;
        pminus *= c;
      }
    }
  } else {
    const double* chi = a.shared_moments ? a.mom + (long)li * a.M
                                         : a.mom + (P.gw * L_ + li) * a.M;
    const double f = (mode == 3) ? a.fdm[P.task] : 0.0;
    const int twoM = 2 * a.n_real;
    int Lm = a.M;
    double inv1mf = 1.0;
    if (mode == 3) {
      inv1mf = 1.0 / (1.0 - f);
      Lm = twoM;
    } else if (mode == 2) {
      Lm = a.M < twoM ? a.M : twoM;
    }
    double sign = 1.0, spj = 0.0, smj = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) Ppp[r] = Ppm[r] = 0.0;
    for (int ll = 0; ll < Lm; ++ll) {
      const double chil = (mode == 3) ? (chi[ll] - f) * inv1mf : chi[ll];
      const double cl = (2 * ll + 1) * chil;
      const double* Pt = a.Pl + ll * N;
      const double plj = Pt[jj];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double term = cl * Pt[ii[r]] * plj;
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
    // Hansen column normalisation over the rows of the lane's layer, in order
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Ppp[r] = inb[r] ? Ppp[r] / (2.0 * kPi) : 0.0;
      Ppm[r] = inb[r] ? Ppm[r] / (2.0 * kPi) : 0.0;
      dtile[j * TS + rows[r]] = inb[r] ? (Ppp[r] + Ppm[r]) * a.wt[ii[r]] : 0.0;
    }
    wave_sync();
    double csum = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) csum += dtile[j * TS + i];
    wave_sync();
    if (csum > 0.0) {
      const double corr = 1.0 / (2.0 * kPi * csum);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Ppp[r] *= corr;
        Ppm[r] *= corr;
      }
    }
    if (a.has_solar) {
      pplus = spj / (2.0 * kPi);
      pminus = smj / (2.0 * kPi);
      const double s = pair_colsum((pplus + pminus) * wt_j, L, pv);
      if (s > 0.0) {
        const double corr = 1.0 / (2.0 * kPi * s);
        pplus *= corr;
        pminus *= corr;
      }
    }
  }
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(omega);
  st.nn = nn0 < 1 ? 1 : nn0;
  st.tau0 = tau * ldexp(1.0, -st.nn);
  const double tau0 = st.tau0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = rows[r];
    const double mui = a.mu[ii[r]];
    const double gpp = (((i == j) ? 1.0 : 0.0) + (-con) * (Ppp[r] * wt_j)) / mui;
    const double gpm = con * (Ppm[r] * wt_j) / mui;
    st.Tv[r] = inb[r] ? ((i == j) ? 1.0 : 0.0) - tau0 * gpp : 0.0;
    st.Sv[r] = inb[r] ? tau0 * gpp : 0.0;
    st.Rv[r] = inb[r] ? tau0 * gpm : 0.0;
  }
  st.y = (1.0 - omega) * tau0 / mu_j;
  st.sus = 0.0;
  st.sds = 0.0;
  st.gam = 0.0;
  if (a.has_solar) {
    const double Ftop = a.F0 * exp(-a.tau_cum[(long)P.w * (L_ + 1) + P.l] / a.mu0);
    const double base = omega * tau0 / mu_j * Ftop;
    st.sus = base * pminus;
    st.sds = base * pplus;
    st.gam = exp(-tau0 / a.mu0);
  }
  st.gk = 0.5 * tau0;
}


// ---------------------------------------------------------------------------
// forward: phase setup + doubling of the two layers of pair p (FP64 records);
// 4 waves per SIMD: 2.78 ms on the 4096 x 60, N = 8 workload against 3.30 ms
// at 5 (20 registers spilled there)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
layer16p_kernel(FwdArgs a) {
  using R = double;
  __shared__ R smem[4 * Lds16::size];
  const int wv = wave_index();
  const long p = (long)blockIdx.x * 4 + wv;
  const int half = a.L / 2;  // pairs per column
  if (p >= (long)a.W * half) return;
  long tA, tB;
  if (!pair_tasks(a, p, tA, tB)) return;  // left to layer16_kernel<false, true>
  const Lane L;
  R* sm = smem + wv * Lds16::size;
  const PairLane P(a, L, tA, tB);
  R* tile = sm + Lds16::tile;
  R* tile2 = sm + Lds16::tile2;
  R* pv = sm + Lds16::pv;
  R* v0 = sm + Lds16::vs;
  R* v1 = v0 + N;
  R* v2 = v1 + N;
  R* v3 = v2 + N;
  R* out = a.lay + P.task * REC16;
  const int L_ = a.L, j = L.j;
  const double tau = a.tau_s[P.task];
  const double omega_in = a.omega_s[P.task];
  const int mode = a.mode[P.task];
  const double Btop = a.B ? a.B[P.gw * (L_ + 1) + lev(a, P.l)] : 0.0;
  const double Bbot = a.B ? a.B[P.gw * (L_ + 1) + lev(a, P.l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (Bbot - Btop) / tau;

  Start16<R> st;
  pair_start(a, L, P, tau, omega_in, mode, tile, pv, st);
  const int nn = st.nn, nnmax = pair_max(nn);
  Mat<R> Rm, Tm;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Tm.v[r] = st.Tv[r];
    Rm.v[r] = st.Rv[r];
  }
  double y = st.y, z = 0.0, sus = st.sus, sds = st.sds, gam = st.gam, gk = st.gk;

  // ---- doubling recursion (src/doubling.hpp:149-209), both layers ----
  for (int k = 0; k < nnmax; ++k) {
    Mat<R> A, X, Y;
    put<R>(tile, L, Rm);
    if (L.h == 0) {
      v0[j] = sds;
      v1[j] = sus;
    }
    wave_sync();
    pmul(tile, Rm, L, Y);  // P = R^2
    double b1 = 0.0, b2 = 0.0;
    if (a.has_solar) {
      double bs[2];
      matvecs<R, 2>(tile, v0, L, bs);
      b1 = bs[0];
      b2 = bs[1];
    }
    wave_sync();
    put<R>(tile, L, Y);
    wave_sync();
    if (inverse_ns<R, false, true>(Y, tile, tile2, A, L)) {  // (I - R^2)^-1
      right_solve_tile<R>(tile, Tm, X, L);                    // or TG by a solve
    } else {
      put<R>(tile2, L, Tm);
      wave_sync();
      pmul(tile2, A, L, X);  // TG = T (I - R^2)^-1
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
    pmul(tile2, Rm, L, Y);  // TGR
    pmul(tile2, Tm, L, A);  // T_new = TG T
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
    pmul_add(tile, Tm, Rm, L, Rm);  // R_new = R + TGR T
    double s2[2];
    matvecs<R, 2>(tile, v0, L, s2);
    wave_sync();
    Tm = A;
    const double znew = (a1 - s2[0]) + z - gk * y;
    const double ynew = a3 + s2[1] + y;
    z = znew;
    y = ynew;
    if (a.has_solar) {
      sus = c1 + sus;
      sds = c2 + gam * sds;
      gam = gam * gam;
    }
    gk = 2.0 * gk;
    if (k == nn - 1) {
      // this lane's layer is complete: its record (src/doubling.hpp:211-229,
      // sources combined), then zero matrices for the steps of the other layer
      pair_gstore(out, L, P, Rm);
      pair_gstore(out + N * N, L, P, Tm);
      pair_vstore(out + 2 * N * N, L, P, (y * B_bar + z * B_d) + sus);
      pair_vstore(out + 2 * N * N + N, L, P, (y * B_bar - z * B_d) + sds);
      if (a.basis) {
        pair_vstore(a.basis + P.task * 2 * N, L, P, y);
        pair_vstore(a.basis + P.task * 2 * N + N, L, P, z);
      }
      if (L.h == 0 && P.jj == 0) a.lflag[P.task] = 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) Rm.v[r] = Tm.v[r] = 0.0;
    }
  }
}



// ---------------------------------------------------------------------------
// reverse: layer16_adjoint_task for the two layers of a pair -- the
// doubling recomputed with per-step checkpoints of both layers in the slab,
// then walked backwards; each layer's steps beyond its own count pass its
// cotangents through unchanged.  Writes the start cotangents of both tasks
// (padded order-16 layout) for start16_adjoint_kernel.
// ---------------------------------------------------------------------------
__device__ void layer16p_adjoint_pair(const FwdArgs& a, const AdjArgs& b, long tA, long tB,
                                      const Lane& L, double* sm, double* slab) {
  using R = double;
  const PairLane P(a, L, tA, tB);
  const int Lay = a.L, j = L.j, jj = P.jj;
  R* t0 = sm + LdsA::t0;
  R* t1 = sm + LdsA::t1;
  R* t2 = sm + LdsA::t2;
  R* t3 = sm + LdsA::t3;
  R* vs = sm + LdsA::vs;
  R* csl = vs + 7 * N;

  const double tau = a.tau_s[P.task];
  const double omega_in = a.omega_s[P.task];
  const int mode = a.mode[P.task];
  const double Btop = a.B ? a.B[P.gw * (Lay + 1) + lev(a, P.l)] : 0.0;
  const double Bbot = a.B ? a.B[P.gw * (Lay + 1) + lev(a, P.l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (Bbot - Btop) / tau;
  // cotangents of the record: R^ = D_l I_dn(l)^T + U_{l+1} I_up(l+1)^T,
  // T^ = D_l I_up(l+1)^T + U_{l+1} I_dn(l)^T over the lane's layer's nodes
  const double* Iw = a.Iv + (long)P.w * (Lay + 1) * 2 * N;
  const double* Aw = a.Av + (long)P.w * (Lay + 1) * 2 * N;
  const double* Dl = Aw + (long)P.l * 2 * N + N;
  const double* U1 = Aw + (long)(P.l + 1) * 2 * N;
  const double idn_j = Iw[(long)P.l * 2 * N + N + jj];
  const double iup_j = Iw[(long)(P.l + 1) * 2 * N + jj];
  const double sub = Dl[jj], sdb = U1[jj];
  Mat<R> Rb, Tb;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = row<R>(L.h, r);
    const bool inb = (i >> 3) == P.hb;
    const int ii = i & 7;
    Rb.v[r] = inb ? Dl[ii] * idn_j + U1[ii] * iup_j : 0.0;
    Tb.v[r] = inb ? Dl[ii] * iup_j + U1[ii] * idn_j : 0.0;
  }

  Mat<R> Rm, Tm;
  double y, z = 0.0, sus, sds, gam, gk, tau0;
  int nn;
  {
    Start16<R> st0;
    pair_start(a, L, P, tau, omega_in, mode, t0, csl, st0);
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
  const int nnmax = pair_max(nn);
  const bool solar = a.has_solar != 0;
  R* s0 = vs;
  R* s1 = vs + N;
  R* s2 = vs + 2 * N;
  R* s3 = vs + 3 * N;
  R* s4 = vs + 4 * N;
  R* s5 = vs + 5 * N;
  R* s6 = vs + 6 * N;

  // ---- doubling of both layers recomputed with checkpoints ----
  for (int k = 0; k < nnmax; ++k) {
    R* st = slab + (long)k * SLAB;
    gstore<R>(st, L, Rm);
    gstore<R>(st + N * N, L, Tm);
    if (L.h == 0) {
      st[3 * N * N + j] = y;
      st[3 * N * N + N + j] = z;
      st[3 * N * N + 2 * N + j] = sus;
      st[3 * N * N + 3 * N + j] = sds;
      if (jj == 0) st[3 * N * N + 4 * N + P.hb] = gam;
    }
    Mat<R> A, Ai, TG, Y, X;
    put<R>(t0, L, Rm);
    if (L.h == 0) {
      s0[j] = sds;
      s1[j] = sus;
    }
    wave_sync();
    pmul(t0, Rm, L, A);  // P = R^2
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
    inverse_ns<R, true, true>(A, t0, t3, Ai, L);  // (I - R^2)^-1
    gstore<R>(st + 2 * N * N, L, Ai);
    put<R>(t1, L, Tm);
    wave_sync();
    pmul(t1, Ai, L, TG);  // TG = T (I - R^2)^-1
    put<R>(t2, L, TG);
    if (L.h == 0) {
      s0[j] = z + gk * y;
      s1[j] = y;
      s2[j] = b1 + gam * sus;
      s3[j] = gam * b2 + sds;
    }
    wave_sync();
    const bool last = k == nnmax - 1;  // the last R_new, T_new are never read back
    pmul(t2, Rm, L, Y);            // TGR
    if (!last) pmul(t2, Tm, L, X);  // T_new = TG T
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
    if (!last) pmul_add(t0, Tm, Rm, L, Rm);  // R_new = R + TGR T
    double s2v[2];
    matvecs<R, 2>(t0, s0, L, s2v);
    wave_sync();
    if (!last) Tm = X;
    // a layer past its own count keeps its final y, z (read by the assembly
    // below) and continues on zero matrices
    const bool live = k < nn;
    const double znew = (a1 - s2v[0]) + z - gk * y;
    const double ynew = a3 + s2v[1] + y;
    z = live ? znew : z;
    y = live ? ynew : y;
    if (solar) {
      sus = c1 + sus;
      sds = c2 + gam * sds;
      gam = gam * gam;
    }
    gk = 2.0 * gk;
    if (k == nn - 1) {
#pragma unroll
      for (int r = 0; r < 4; ++r) Rm.v[r] = Tm.v[r] = 0.0;
    }
  }

  // ---- layer assembly, reversed (per layer of the pair) ----
  double yb = (sub + sdb) * B_bar, zb = (sub - sdb) * B_d, susb = sub, sdsb = sdb;
  double g_tau, g_Bt, g_Bb;
  {
    const double gbar = pair_colsum((sub + sdb) * y, L, csl);
    const double gd = pair_colsum((sub - sdb) * z, L, csl);
    g_Bt = 0.5 * gbar - gd / tau;
    g_Bb = 0.5 * gbar + gd / tau;
    g_tau = -gd * B_d / tau;
  }
  double gamb = 0.0, gb = 0.0;
  Mat<R> Rbar = Rb, Tbar = Tb;

  // ---- doubling steps in reverse; a layer past its count passes through ----
  for (int k = nnmax - 1; k >= 0; --k) {
    const R* st = slab + (long)k * SLAB;
    const bool live = k < nn;
    Mat<R> Rk, Tk, Ai, TG, TGR, Tt, Rt;
    gload<R>(st, L, Rk);
    gload<R>(st + N * N, L, Tk);
    gload<R>(st + 2 * N * N, L, Ai);
    const double yk = st[3 * N * N + j], zk = st[3 * N * N + N + j];
    const double susk = st[3 * N * N + 2 * N + j], sdsk = st[3 * N * N + 3 * N + j];
    const double gm = st[3 * N * N + 4 * N + P.hb];
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
    pmul(t0, Ai, L, TG);
    get_t<R>(t0, L, Tt);
    get_t<R>(t2, L, Rt);
    wave_sync();
    put<R>(t1, L, TG);
    wave_sync();
    pmul(t1, Rk, L, TGR);
    const double u = zk + g * yk;
    const double x_zb = tmatvec<R>(TG, s0, L);
    const double x_yb = tmatvec<R>(TG, s1, L);
    const double r_zb = tmatvec<R>(TGR, s0, L);
    const double r_yb = tmatvec<R>(TGR, s1, L);
    const double ub = x_zb - r_zb;
    const double yb_n = ((yb - g * zb) + (x_yb + r_yb)) + g * ub;
    const double zb_n = zb + ub;
    const double gb_n = 2.0 * gb + (ub - zb) * yk;
    double v2 = 0.0, v3 = 0.0, b1b = 0.0, b2b = 0.0, susb_n = susb, sdsb_n = sdsb,
           gamb_n = gamb;
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
        s0[j] = b2b;
      }
      wave_sync();
      const double rt_b1 = tmatvec<R>(Rk, s6, L);
      const double rt_b2 = tmatvec<R>(Rk, s0, L);
      susb_n = (susb + gm * v2b) + rt_b2;
      sdsb_n = (gm * sdsb + v3b) + rt_b1;
      gamb_n = 2.0 * gm * gamb + ((sdsb * sdsk + v2b * susk) + v3b * bs[1]);
    }
    // matrices, reversed
    d4 acc = zero4<R>();
    put<R>(t0, L, Rbar);
    put<R>(t3, L, Tbar);
    wave_sync();
    outer4<R>(acc, -zb, yb, 0.0, 0.0, u, yk, 0.0, 0.0, L);
    pmac(acc, t0, Tt, L);  // TGR^ = R'^ T^T + vector terms
    pmask(acc, L);
    const Mat<R> TGRb = mat_of<R>(acc);
    acc = zero4<R>();
    outer4<R>(acc, zb, yb, susb, sdsb, u, yk, v2, v3, L);
    pmac(acc, t3, Tt, L);  // TG^ = T'^ T^T + ...
    wave_sync();
    put<R>(t0, L, TGRb);
    wave_sync();
    pmac(acc, t0, Rt, L);  // ... + TGR^ R^T
    pmask(acc, L);
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
    pmac(acc, t3, AiT, L);  // W = TG^ (I - R^2)^-T
    pmask(acc, L);
    const Mat<R> W = mat_of<R>(acc);
    acc = zero4<R>();
    pmac_t(acc, t1, W, L);  // Abn = -A^ = TG^T W
    pmask(acc, L);
    const Mat<R> Abn = mat_of<R>(acc);
    acc = acc_of<R>(Rbar);
    outer4<R>(acc, b1b, b2b, 0.0, 0.0, sdsk, susk, 0.0, 0.0, L);
    pmac_t(acc, t1, TGRb, L);  // + TG^T TGR^
    wave_sync();
    put<R>(t0, L, Abn);
    wave_sync();
    pmac(acc, t0, Rt, L);     // + Abn R^T
    pmac_t(acc, t2, Abn, L);  // + R^T Abn
    pmask(acc, L);
    const Mat<R> Rbar_n = mat_of<R>(acc);
    wave_sync();
    put<R>(t3, L, TGR);
    wave_sync();
    acc = acc_of<R>(W);
    pmac_t(acc, t3, Rbar, L);  // + TGR^T R'^
    pmac_t(acc, t1, Tbar, L);  // + TG^T T'^
    pmask(acc, L);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Tbar.v[r] = live ? acc[r] : Tbar.v[r];
      Rbar.v[r] = live ? Rbar_n.v[r] : Rbar.v[r];
    }
    yb = live ? yb_n : yb;
    zb = live ? zb_n : zb;
    susb = live ? susb_n : susb;
    sdsb = live ? sdsb_n : sdsb;
    gb = live ? gb_n : gb;
    gamb = live ? gamb_n : gamb;
    wave_sync();
  }

  // ---- cotangents of both doubling starts, for start16_adjoint_kernel ----
  const double g_sum = pair_colsum(gb, L, csl);
  const double gam_sum = solar ? pair_colsum(gamb, L, csl) : 0.0;
  double* sc = b.scot + P.task * SCOT;
  pair_gstore(sc, L, P, Rbar);
  pair_gstore(sc + N * N, L, P, Tbar);
  pair_vstore(sc + 2 * N * N, L, P, yb);
  pair_vstore(sc + 2 * N * N + N, L, P, susb);
  pair_vstore(sc + 2 * N * N + 2 * N, L, P, sdsb);
  if (L.h == 0 && jj == 0) {
    sc[2 * N * N + 3 * N] = g_sum;
    sc[2 * N * N + 3 * N + 1] = gam_sum;
    b.g_taus[P.task] = g_tau;  // completed by start16_adjoint_kernel
    b.g_Bt[P.task] = g_Bt;
    b.g_Bb[P.task] = g_Bb;
  }
}

// persistent grid over the pairs; 2 waves per SIMD (256 registers: the pair
// state spills at 3 -- 7.2 against 7.8 ms on the 4096 x 60, N = 8 workload)
constexpr int PAIR_REV_WAVES_PER_SIMD = 2;
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PAIR_REV_WAVES_PER_SIMD, 8)))
layer16p_adjoint_kernel(FwdArgs a, AdjArgs b) {
  __shared__ double smem[4 * LdsA::size];
  const int wv = wave_index();
  const int gwave = blockIdx.x * 4 + wv;
  const Lane L;
  double* sm = smem + wv * LdsA::size;
  double* slab = b.slab + (long)gwave * slab_stride(b.nn_cap);
  const long npair = (long)a.W * (a.L / 2);
  for_tasks<TASK_CHUNK>(b.queue, npair, [&](long p) {
    long tA, tB;
    if (!pair_tasks(a, p, tA, tB)) return;  // left to layer16_adjoint_kernel<false, true>
    if (doubling_count(a.tau_s[tA], a.omega_s[tA]) > b.nn_cap ||
        doubling_count(a.tau_s[tB], a.omega_s[tB]) > b.nn_cap) {
      // beyond the slab: flagged, no gradients (as layer16_adjoint_task)
      if (L.l == 0) atomicOr(b.status, 1);
      for (long t = tA; ; t = tB) {
        const int l = (int)(t % a.L);
        const long gw = (long)(a.w0 + t / a.L);
        double* chi_out = a.shared_moments
                              ? b.g_chi + t * a.M
                              : (b.d_mom ? b.d_mom + (gw * a.L + lin(a, l)) * a.M : nullptr);
        if (chi_out)
          for (int m = L.l; m < a.M; m += 64) chi_out[m] = 0.0;
        if (L.l == 0) {
          b.g_taus[t] = b.g_omgs[t] = b.g_Bt[t] = b.g_Bb[t] = b.g_Ftop[t] = b.g_f[t] = 0.0;
        }
        if (t == tB) break;
      }
      return;
    }
    layer16p_adjoint_pair(a, b, tA, tB, L, sm, slab);
  });
}

}  // namespace w16
}  // namespace dev
}  // namespace adrt
