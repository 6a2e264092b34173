// adrt_adj32.hpp -- reverse pass (adjoint) of the N = 32 solve.
//
// The N = 16 reverse pass (adrt_adj16.hpp) on the N = 32 layouts of
// adrt_wave32.hpp (four 16 x 16 blocks per wave, rows 16 (b & 1) + h + 4 r of
// column 16 (b >> 1) + j; vectors as the pair of elements j, 16 + j):
//
//  1. adding32_kernel<true> and, for the columns the prep kernel flags,
//     adding32_composite_kernel<true> solve the transposed interface system
//     for the adjoint vectors D_l, U_l;
//  2. layer32_adjoint_kernel, one wave per (wavenumber, layer) on a
//     persistent grid: the cotangents of the layer record are the outer
//     products of forward and adjoint interface vectors,
//        R^ = D_l I_dn(l)^T + U_{l+1} I_up(l+1)^T,
//        T^ = D_l I_up(l+1)^T + U_{l+1} I_dn(l)^T,  s_up^ = D_l, s_dn^ = U_{l+1};
//     the doubling (src/doubling.hpp:149-209) is recomputed from the
//     thin-layer start with per-step checkpoints of R, T and (I - R^2)^-1 in
//     the wave's slab, then walked backwards with matrix-core products
//     (the step's reverse is the N = 16 one, adrt_adj16.hpp);
//  3. start32_adjoint_kernel: the thin-layer start, the Hansen normalisation
//     and the Legendre sums (src/doubling.hpp:83-147,
//     src/phase_matrix.hpp:61-181) reversed;
//  4. column32_adjoint_kernel: the column terms (w16::column_adjoint).
#pragma once

#include "adrt_adj16.hpp"
#include "adrt_wave32.hpp"

namespace adrt {
namespace dev {
namespace w32 {

using w16::AdjArgs;
using w16::col_reduce;
using w16::wave_sum;

// per-step checkpoint: R, T, (I - R^2)^-1 (column-major); y, z, s_up, s_dn;
// gamma
constexpr int SLAB32 = 3 * N * N + 4 * N + 2;
constexpr int SL_Y = 3 * N * N, SL_Z = SL_Y + N, SL_SU = SL_Z + N, SL_SD = SL_SU + N,
              SL_GAM = SL_SD + N;
static_assert(SL_GAM < SLAB32, "checkpoint fields past the slab stride");
__host__ __device__ constexpr long slab_stride32(int nn_cap) { return (long)nn_cap * SLAB32; }

// cotangents of the doubling start: R0, T0 (column-major); y0, s_up0, s_dn0;
// column sums of the g and gamma cotangents
constexpr int SCOT32 = 2 * N * N + 3 * N + 2;

// one wave per workgroup: four 32 x 33 tiles, seven vector slots and a
// column-sum slot (35.8 KB; four workgroups per CU)
struct LdsA32 {
  static constexpr int t0 = 0, t1 = TILE, t2 = 2 * TILE, t3 = 3 * TILE, vs = 4 * TILE;
  static constexpr int csl = vs + 7 * N, size = csl + N;
};
static_assert(LdsA32::t3 + TILE <= LdsA32::vs && LdsA32::vs + 7 * N <= LdsA32::csl &&
                  LdsA32::csl + N <= LdsA32::size && LdsA32::size * 8 <= 40 * 1024,
              "N = 32 reverse per-wave LDS regions overlap");

// C += X Y (X in a tile, Y in registers)
__device__ __forceinline__ void mac(Mat& C, const double* X, const Mat& Y, const Lane& L) {
  mul_impl<true>(X, Y, C, L, C);
}

// C += X^T Y (X in a tile, read transposed; Y in registers)
__device__ __forceinline__ void mac_t(Mat& C, const double* X, const Mat& Y, const Lane& L) {
#pragma unroll
  for (int J = 0; J < 2; ++J) {
    fence_sched();
    d4 acc[2];
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      const int b = I + 2 * J;
      acc[I] = d4{C.v[b][0], C.v[b][1], C.v[b][2], C.v[b][3]};
    }
#pragma unroll
    for (int K = 0; K < 2; ++K)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double y = Y.v[K + 2 * J][r];
#pragma unroll
        for (int I = 0; I < 2; ++I)  // X^T(16 I + j, 16 K + h + 4 r) = X(16 K + h + 4 r, 16 I + j)
          acc[I] = mfma(X[(16 * I + L.j) * TS + 16 * K + rw(L.h, r)], y, acc[I]);
      }
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) C.v[I + 2 * J][r] = acc[I][r];
    fence_sched();
  }
}

// C += a0 b0^T + a1 b1^T + a2 b2^T + a3 b3^T (one K = 4 instruction per
// block: lane group h supplies the pair (a_h, b_h))
__device__ __forceinline__ void outer4(Mat& C, const Vec& a0, const Vec& a1, const Vec& a2,
                                       const Vec& a3, const Vec& b0, const Vec& b1,
                                       const Vec& b2, const Vec& b3, const Lane& L) {
  const Vec av = L.h == 0 ? a0 : L.h == 1 ? a1 : L.h == 2 ? a2 : a3;
  const Vec bv = L.h == 0 ? b0 : L.h == 1 ? b1 : L.h == 2 ? b2 : b3;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    d4 acc{C.v[b][0], C.v[b][1], C.v[b][2], C.v[b][3]};
    acc = mfma((b & 1) ? av.hi : av.lo, (b >> 1) ? bv.hi : bv.lo, acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) C.v[b][r] = acc[r];
  }
}

// M^T v for M in registers and v in an LDS slot of 32 values
__device__ __forceinline__ Vec tmatvec(const Mat& M, const double* v, const Lane& L) {
  double p[2] = {0.0, 0.0};
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) p[b >> 1] = fma_(M.v[b][r], v[bi(b) + rw(L.h, r)], p[b >> 1]);
  return Vec{col_reduce<double>(p[0], L), col_reduce<double>(p[1], L)};
}

// ordered sum over the 32 columns (one Vec per lane column), as the
// reference's quadrature loops
__device__ __forceinline__ double colsum32(const Vec& v, const Lane& L, double* slot) {
  put_vec(slot, L, v);
  wave_sync();
  double s = 0.0;
#pragma unroll 4
  for (int i = 0; i < N; ++i) s += slot[i];
  wave_sync();
  return s;
}

__device__ __forceinline__ Vec vsub(const Vec& a, const Vec& b) { return Vec{a.lo - b.lo, a.hi - b.hi}; }
__device__ __forceinline__ Vec vscale(double s, const Vec& a) { return Vec{s * a.lo, s * a.hi}; }

// reverse of a transparent (no dependence) or purely absorbing layer
// (src/doubling.hpp:52-77); returns false for a scattering layer
__device__ __forceinline__ bool nonscat32_reverse(const FwdArgs& a, const Lane& L, const Vec& Dl,
                                                  const Vec& U1, const Vec& idn, const Vec& iup,
                                                  double tau, double omega_in, double B_bar,
                                                  double B_d, double* csl, double& g_tau,
                                                  double& g_Bt, double& g_Bb) {
  if (tau <= 0.0) return true;
  if (omega_in > 0.0) return false;
  const double mu_c[2] = {a.mu[L.j], a.mu[16 + L.j]};
  const double dl[2] = {Dl.lo, Dl.hi}, u1[2] = {U1.lo, U1.hi};
  const double id[2] = {idn.lo, idn.hi}, iu[2] = {iup.lo, iup.hi};
  double gt[2], gb[2], gdd[2];
  const double dBd = -B_d / tau;
#pragma unroll
  for (int J = 0; J < 2; ++J) {
    const double tex = -tau / mu_c[J];
    const bool live = tex > -200.0;
    const double t = live ? exp(tex) : 0.0;
    const double dt = live ? -t / mu_c[J] : 0.0;
    const double omt = 1.0 - t;
    const double slope = mu_c[J] * omt - 0.5 * tau * (1.0 + t);
    const double dslope = -mu_c[J] * dt - 0.5 * (1.0 + t) - 0.5 * tau * dt;
    const double tbc = dl[J] * iu[J] + u1[J] * id[J];  // T^(c, c)
    const double dsu = -B_bar * dt + dBd * slope + B_d * dslope;
    const double dsd = -B_bar * dt - dBd * slope - B_d * dslope;
    gt[J] = tbc * dt + dl[J] * dsu + u1[J] * dsd;
    gb[J] = (dl[J] + u1[J]) * omt;
    gdd[J] = (dl[J] - u1[J]) * slope;
  }
  g_tau = colsum32(Vec{gt[0], gt[1]}, L, csl);
  const double gbar = colsum32(Vec{gb[0], gb[1]}, L, csl);
  const double gd = colsum32(Vec{gdd[0], gdd[1]}, L, csl);
  g_Bt = 0.5 * gbar - gd / tau;
  g_Bb = 0.5 * gbar + gd / tau;
  return true;
}

__device__ void layer32_adjoint_task(const FwdArgs& a, const AdjArgs& b, long task, const Lane& L,
                                     double* sm, double* slab) {
  const int Lay = a.L;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  double* t0 = sm + LdsA32::t0;
  double* t1 = sm + LdsA32::t1;
  double* t2 = sm + LdsA32::t2;
  double* t3 = sm + LdsA32::t3;
  double* s0 = sm + LdsA32::vs;
  double* s1 = s0 + N;
  double* s2 = s0 + 2 * N;
  double* s3 = s0 + 3 * N;
  double* s4 = s0 + 4 * N;
  double* s6 = s0 + 6 * N;
  double* csl = sm + LdsA32::csl;

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
  // the layer's adjoint and forward interface vectors
  const double* Iw = a.Iv + (long)w * (Lay + 1) * 2 * N;
  const double* Aw = a.Av + (long)w * (Lay + 1) * 2 * N;
  const double* Dlp = Aw + (long)l * 2 * N + N;    // D_l
  const double* U1p = Aw + (long)(l + 1) * 2 * N;  // U_{l+1}
  const Vec Dl = gvec(Dlp, L), U1 = gvec(U1p, L);
  const Vec idn = gvec(Iw + (long)l * 2 * N + N, L);    // I_dn(l)
  const Vec iup = gvec(Iw + (long)(l + 1) * 2 * N, L);  // I_up(l+1)
  double g_tau = 0.0, g_Bt = 0.0, g_Bb = 0.0;
  auto finish = [&]() {
    if (chi_out)
      for (int m = L.l; m < a.M; m += 64) chi_out[m] = 0.0;
    if (L.l == 0) {
      b.g_taus[task] = g_tau;
      b.g_omgs[task] = 0.0;
      b.g_Bt[task] = g_Bt;
      b.g_Bb[task] = g_Bb;
      b.g_Ftop[task] = 0.0;
      b.g_f[task] = 0.0;
    }
  };
  if (nonscat32_reverse(a, L, Dl, U1, idn, iup, tau, omega_in, B_bar, B_d, csl, g_tau, g_Bt,
                        g_Bb)) {
    finish();
    return;
  }

  // the thin-layer start as the forward computed it; its reverse (phase
  // matrices, Hansen normalisation, Legendre sums) is start32_adjoint_kernel's
  Mat Rm, Tm;
  Vec y, z{0.0, 0.0}, sus, sds;
  double gam, gk, tau0;
  int nn;
  {
    Start32 st;
    start32<double, false>(a, task, w, l, gw, L, tau, omega_in, mode, st);
    Rm = st.R;
    Tm = st.T;
    y = st.y;
    sus = st.sus;
    sds = st.sds;
    gam = st.gam;
    gk = st.gk;
    tau0 = st.tau0;
    nn = st.nn;
  }
  if (nn > b.nn_cap) {
    if (L.l == 0) atomicOr(b.status, 1);
    finish();
    return;
  }
  const bool solar = a.has_solar != 0;

  // ---- doubling recomputed with checkpoints (as layer32_kernel) ----
  for (int k = 0; k < nn; ++k) {
    double* st = slab + (long)k * SLAB32;
    gstore(st, L, Rm);
    gstore(st + N * N, L, Tm);
    store_vec(st + SL_Y, L, y);
    store_vec(st + SL_Z, L, z);
    store_vec(st + SL_SU, L, sus);
    store_vec(st + SL_SD, L, sds);
    if (L.l == 0) st[SL_GAM] = gam;
    put(t0, L, Rm);
    put_vec(s0, L, sds);
    put_vec(s1, L, sus);
    wave_sync();
    Mat P;
    mul(t0, Rm, L, P);  // R^2
    Vec b12[2] = {Vec{0.0, 0.0}, Vec{0.0, 0.0}};
    if (solar) matvecs<2>(t0, s0, L, b12);  // R s_dn, R s_up
    wave_sync();
    put(t0, L, P);
    wave_sync();
    Mat X;
    if (inverse_ns(P, t0, t3, X, L)) {  // (I - R^2)^-1, or by LU where NS does not apply
      Mat I1;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) I1.v[q][r] = delta(L, q, r);
      right_solve_tile(t0, t3, I1, X, L);
    }
    gstore(st + 2 * N * N, L, X);
    put(t1, L, Tm);
    wave_sync();
    Mat TG;
    mul(t1, X, L, TG);  // TG = T (I - R^2)^-1
    put(t2, L, TG);
    put_vec(s0, L, vadd(z, vscale(gk, y)));
    put_vec(s1, L, y);
    put_vec(s2, L, vadd(b12[0], vscale(gam, sus)));
    put_vec(s3, L, vadd(vscale(gam, b12[1]), sds));
    wave_sync();
    const bool last = k == nn - 1;  // the last R', T' are never read back
    Mat Y;
    mul(t2, Rm, L, Y);  // TGR
    if (!last) mul(t2, Tm, L, X);  // T' = TG T
    Vec s4v[4];
    if (solar) {
      matvecs<4>(t2, s0, L, s4v);
    } else {
      Vec s2v[2];
      matvecs<2>(t2, s0, L, s2v);
      s4v[0] = s2v[0];
      s4v[1] = s2v[1];
      s4v[2] = Vec{0.0, 0.0};
      s4v[3] = Vec{0.0, 0.0};
    }
    wave_sync();
    put(t0, L, Y);
    wave_sync();
    if (!last) mul_add(t0, Tm, Rm, L, Rm);  // R' = R + TGR T
    Vec s2v[2];
    matvecs<2>(t0, s0, L, s2v);  // TGR (z + g y), TGR y
    wave_sync();
    if (!last) Tm = X;
    const Vec znew{(s4v[0].lo - s2v[0].lo) + z.lo - gk * y.lo,
                   (s4v[0].hi - s2v[0].hi) + z.hi - gk * y.hi};
    const Vec ynew{s4v[1].lo + s2v[1].lo + y.lo, s4v[1].hi + s2v[1].hi + y.hi};
    z = znew;
    y = ynew;
    if (solar) {
      sus = vadd(s4v[2], sus);
      sds = vadd(s4v[3], vscale(gam, sds));
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // ---- layer assembly, reversed: s_up = y B_bar + z B_d + s_up_sol, s_dn = y B_bar - z B_d + s_dn_sol
  const Vec sub = Dl, sdb = U1;
  Vec yb = vscale(B_bar, vadd(sub, sdb)), zb = vscale(B_d, vsub(sub, sdb));
  Vec susb = sub, sdsb = sdb;
  {
    const double gbar = colsum32(vmul(vadd(sub, sdb), y), L, csl);
    const double gd = colsum32(vmul(vsub(sub, sdb), z), L, csl);
    g_Bt = 0.5 * gbar - gd / tau;
    g_Bb = 0.5 * gbar + gd / tau;
    g_tau = -gd * B_d / tau;
  }
  Vec gamb{0.0, 0.0}, gb{0.0, 0.0};  // per-column parts of the scalar cotangents
  Mat Rbar, Tbar;
  {
    double dli[2][4], u1i[2][4];
    rows_of(Dlp, L, dli);
    rows_of(U1p, L, u1i);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double d = dli[q & 1][r], u = u1i[q & 1][r];
        Rbar.v[q][r] = d * col_of(idn, q) + u * col_of(iup, q);
        Tbar.v[q][r] = d * col_of(iup, q) + u * col_of(idn, q);
      }
  }

  // ---- doubling steps in reverse (the N = 16 step, adrt_adj16.hpp) ----
  for (int k = nn - 1; k >= 0; --k) {
    const double* st = slab + (long)k * SLAB32;
    Mat Rk, Tk, Ai, TG, TGR, Tt, Rt;
    gload(st, L, Rk);
    gload(st + N * N, L, Tk);
    gload(st + 2 * N * N, L, Ai);
    const Vec yk = gvec(st + SL_Y, L), zk = gvec(st + SL_Z, L);
    const Vec susk = gvec(st + SL_SU, L), sdsk = gvec(st + SL_SD, L);
    const double gm = st[SL_GAM];
    const double g = 0.5 * tau0 * ldexp(1.0, k);
    put(t0, L, Tk);
    put(t2, L, Rk);
    put_vec(s0, L, zb);
    put_vec(s1, L, yb);
    put_vec(s2, L, susb);
    put_vec(s3, L, sdsb);
    put_vec(s4, L, sdsk);
    put_vec(s4 + N, L, susk);
    wave_sync();
    mul(t0, Ai, L, TG);
    get_t(t0, L, Tt);
    get_t(t2, L, Rt);
    wave_sync();
    put(t1, L, TG);
    wave_sync();
    mul(t1, Rk, L, TGR);
    const Vec u = vadd(zk, vscale(g, yk));
    // vector recursions, reversed
    const Vec x_zb = tmatvec(TG, s0, L), x_yb = tmatvec(TG, s1, L);
    const Vec r_zb = tmatvec(TGR, s0, L), r_yb = tmatvec(TGR, s1, L);
    const Vec ub = vsub(x_zb, r_zb);
    const Vec yb_n{((yb.lo - g * zb.lo) + (x_yb.lo + r_yb.lo)) + g * ub.lo,
                   ((yb.hi - g * zb.hi) + (x_yb.hi + r_yb.hi)) + g * ub.hi};
    const Vec zb_n = vadd(zb, ub);
    gb = Vec{2.0 * gb.lo + (ub.lo - zb.lo) * yk.lo, 2.0 * gb.hi + (ub.hi - zb.hi) * yk.hi};
    Vec v2{0.0, 0.0}, v3{0.0, 0.0}, b1b{0.0, 0.0}, b2b{0.0, 0.0};
    Vec susb_n = susb, sdsb_n = sdsb;
    if (solar) {
      const Vec v2b = tmatvec(TG, s2, L), v3b = tmatvec(TG, s3, L);
      Vec bs[2];
      matvecs<2>(t2, s4, L, bs);  // R s_dn, R s_up
      v2 = vadd(bs[0], vscale(gm, susk));
      v3 = vadd(vscale(gm, bs[1]), sdsk);
      b1b = v2b;
      b2b = vscale(gm, v3b);
      wave_sync();
      put_vec(s6, L, b1b);
      put_vec(s0, L, b2b);  // s0 is free again (read above)
      wave_sync();
      const Vec rt_b1 = tmatvec(Rk, s6, L), rt_b2 = tmatvec(Rk, s0, L);
      susb_n = vadd(vadd(susb, vscale(gm, v2b)), rt_b2);
      sdsb_n = vadd(vadd(vscale(gm, sdsb), v3b), rt_b1);
      gamb = Vec{2.0 * gm * gamb.lo + ((sdsb.lo * sdsk.lo + v2b.lo * susk.lo) + v3b.lo * bs[1].lo),
                 2.0 * gm * gamb.hi + ((sdsb.hi * sdsk.hi + v2b.hi * susk.hi) + v3b.hi * bs[1].hi)};
    }
    // matrices, reversed
    const Vec zero{0.0, 0.0};
    put(t0, L, Rbar);
    put(t3, L, Tbar);
    wave_sync();
    Mat TGRb;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) TGRb.v[q][r] = 0.0;
    outer4(TGRb, vscale(-1.0, zb), yb, zero, zero, u, yk, zero, zero, L);
    mac(TGRb, t0, Tt, L);  // TGR^ = R'^ T^T + vector terms
    Mat TGb;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) TGb.v[q][r] = 0.0;
    outer4(TGb, zb, yb, susb, sdsb, u, yk, v2, v3, L);
    mac(TGb, t3, Tt, L);  // TG^ = T'^ T^T + ...
    wave_sync();
    put(t0, L, TGRb);
    wave_sync();
    mac(TGb, t0, Rt, L);  // ... + TGR^ R^T
    wave_sync();
    put(t3, L, Ai);
    wave_sync();
    Mat AiT;
    get_t(t3, L, AiT);
    wave_sync();
    put(t3, L, TGb);
    wave_sync();
    Mat W;
    mul(t3, AiT, L, W);  // W = TG^ (I - R^2)^-T
    Mat Abn;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) Abn.v[q][r] = 0.0;
    mac_t(Abn, t1, W, L);  // Abn = -A^ = TG^T W
    Mat Rn = Rbar;
    outer4(Rn, b1b, b2b, zero, zero, sdsk, susk, zero, zero, L);
    mac_t(Rn, t1, TGRb, L);  // + TG^T TGR^
    wave_sync();
    put(t0, L, Abn);
    wave_sync();
    mac(Rn, t0, Rt, L);     // + Abn R^T
    mac_t(Rn, t2, Abn, L);  // + R^T Abn
    wave_sync();
    put(t3, L, TGR);
    wave_sync();
    Mat Tn = W;
    mac_t(Tn, t3, Rbar, L);  // + TGR^T R'^
    mac_t(Tn, t1, Tbar, L);  // + TG^T T'^
    Tbar = Tn;
    Rbar = Rn;
    yb = yb_n;
    zb = zb_n;
    susb = susb_n;
    sdsb = sdsb_n;
    wave_sync();
  }

  // ---- cotangents of the doubling start, for start32_adjoint_kernel ----
  const double g_sum = colsum32(gb, L, csl);
  const double gam_sum = solar ? colsum32(gamb, L, csl) : 0.0;
  double* sc = b.scot + task * SCOT32;
  gstore(sc, L, Rbar);
  gstore(sc + N * N, L, Tbar);
  store_vec(sc + 2 * N * N, L, yb);
  store_vec(sc + 2 * N * N + N, L, susb);
  store_vec(sc + 2 * N * N + 2 * N, L, sdsb);
  if (L.l == 0) {
    sc[2 * N * N + 3 * N] = g_sum;
    sc[2 * N * N + 3 * N + 1] = gam_sum;
    b.g_taus[task] = g_tau;  // completed by start32_adjoint_kernel
    b.g_Bt[task] = g_Bt;
    b.g_Bb[task] = g_Bb;
  }
}

// one wave per workgroup, 1 wave per SIMD (the reverse step keeps about ten
// 32 x 32 matrices in registers); a persistent grid, one slab per wave
constexpr int REV32_WAVES_PER_SIMD = 1;

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(REV32_WAVES_PER_SIMD, 8)))
layer32_adjoint_kernel(FwdArgs a, AdjArgs b) {
  __shared__ double smem[LdsA32::size];
  const int gwave = blockIdx.x;
  const Lane L;
  double* slab = b.slab + (long)gwave * slab_stride32(b.nn_cap);
  const long ntask = (long)a.W * a.L;
  w16::for_tasks<2>(b.queue, ntask,
                     [&](long task) { layer32_adjoint_task(a, b, task, L, smem, slab); });
}

// ---------------------------------------------------------------------------
// The thin-layer start, the Hansen normalisation and the Legendre sums of
// each scattering layer, reversed, from the start cotangents the doubling
// reverse left in b.scot; one wave per (wavenumber, layer).  The phase
// matrices are recomputed as start32 forms them (src/phase_matrix.hpp:61-181),
// keeping the sums before the normalisation.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8)))
start32_adjoint_kernel(FwdArgs a, AdjArgs b) {
  const long task = blockIdx.x;
  const int Lay = a.L;
  if (task >= (long)a.W * Lay) return;
  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  if (tau <= 0.0 || omega_in <= 0.0) return;  // finished by the doubling reverse
  const Lane L;
  const int j = L.j, h = L.h;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  const int mode = a.mode[task];
  const double omega = omega_in > 1.0 ? 1.0 : omega_in;
  const double con = 2.0 * omega * kPi;
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(omega);
  const int nn = nn0 < 1 ? 1 : nn0;
  if (nn > b.nn_cap) return;  // reported through b.status
  const double tau0 = tau * ldexp(1.0, -nn);
  const int li = lin(a, l);
  const double* chi = a.shared_moments ? a.mom + (long)li * a.M : a.mom + (gw * Lay + li) * a.M;
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
  const bool solar = a.has_solar != 0;
  const double mu_c[2] = {a.mu[j], a.mu[16 + j]};
  const double wt_c[2] = {a.wt[j], a.wt[16 + j]};

  // phase sums / 2 pi (Q, before the Hansen normalisation)
  Mat Qpp, Qpm;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Qpp.v[q][r] = 0.0;
      Qpm.v[q][r] = 0.0;
    }
  double sp[2] = {0.0, 0.0}, smv[2] = {0.0, 0.0};
  double sign = 1.0;
  for (int ll = 0; ll < Lm; ++ll) {
    const double chil = (mode == 3) ? (chi[ll] - f) * inv1mf : chi[ll];
    const double cl = (2 * ll + 1) * chil;
    const double* Pl = a.Pl + ll * N;
    const double plc[2] = {Pl[j], Pl[16 + j]};
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double pi = cl * Pl[16 * I + rw(h, r)];
#pragma unroll
        for (int J = 0; J < 2; ++J) {
          const double term = pi * plc[J];
          Qpp.v[I + 2 * J][r] += term;
          Qpm.v[I + 2 * J][r] += sign * term;
        }
      }
    if (solar) {
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        const double term = cl * plc[J] * a.Pl0[ll];
        sp[J] += term;
        smv[J] += sign * term;
      }
    }
    sign = -sign;
  }
  double csum[2], corr[2];
  {
    double part[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Qpp.v[q][r] = Qpp.v[q][r] / (2.0 * kPi);
        Qpm.v[q][r] = Qpm.v[q][r] / (2.0 * kPi);
        part[q >> 1] += (Qpp.v[q][r] + Qpm.v[q][r]) * a.wt[bi(q) + rw(h, r)];
      }
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      csum[J] = col_reduce<double>(part[J], L);
      corr[J] = csum[J] > 0.0 ? 1.0 / (2.0 * kPi * csum[J]) : 1.0;
    }
  }
  double qplus[2] = {0.0, 0.0}, qminus[2] = {0.0, 0.0}, pplus[2] = {0.0, 0.0},
         pminus[2] = {0.0, 0.0};
  double ssol = 0.0, corr_s = 1.0;
  if (solar) {
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      qplus[J] = sp[J] / (2.0 * kPi);
      qminus[J] = smv[J] / (2.0 * kPi);
    }
    ssol = row_sum<double>((qplus[0] + qminus[0]) * wt_c[0] + (qplus[1] + qminus[1]) * wt_c[1]);
    if (ssol > 0.0) corr_s = 1.0 / (2.0 * kPi * ssol);
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      pplus[J] = ssol > 0.0 ? qplus[J] * corr_s : qplus[J];
      pminus[J] = ssol > 0.0 ? qminus[J] * corr_s : qminus[J];
    }
  }
  double Ftop = 0.0, gam0 = 0.0;
  if (solar) {
    Ftop = a.F0 * exp(-a.tau_cum[(long)w * (Lay + 1) + l] / a.mu0);
    gam0 = exp(-tau0 / a.mu0);
  }

  // start cotangents
  const double* sc = b.scot + task * SCOT32;
  Mat Rb0, Tb0;
  gload(sc, L, Rb0);
  gload(sc + N * N, L, Tb0);
  const Vec yb = gvec(sc + 2 * N * N, L), susb = gvec(sc + 2 * N * N + N, L);
  const Vec sdsb = gvec(sc + 2 * N * N + 2 * N, L);
  const double g_sum = sc[2 * N * N + 3 * N], gam_sum = sc[2 * N * N + 3 * N + 1];

  // thin-layer start reversed (src/doubling.hpp:83-147)
  double tb_l = 0.0, cb_l = 0.0, ob_l = 0.0, fb_l = 0.0;
  Mat Pb_pp, Pb_pm;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = bi(q) + rw(h, r);
      const double mui = a.mu[i];
      const double wc = wt_c[q >> 1];
      const double dij = (i == bj(q) + j) ? 1.0 : 0.0;
      const double Ppp = csum[q >> 1] > 0.0 ? Qpp.v[q][r] * corr[q >> 1] : Qpp.v[q][r];
      const double Ppm = csum[q >> 1] > 0.0 ? Qpm.v[q][r] * corr[q >> 1] : Qpm.v[q][r];
      const double tb = Tb0.v[q][r], rb = Rb0.v[q][r];
      tb_l += tb * (-(dij - con * Ppp * wc) / mui) + rb * (con * Ppm * wc / mui);
      cb_l += (tb * Ppp + rb * Ppm) * (tau0 * wc / mui);
      Pb_pp.v[q][r] = tb * (tau0 * con * wc / mui);
      Pb_pm.v[q][r] = rb * (tau0 * con * wc / mui);
    }
  double pplusb[2] = {0.0, 0.0}, pminusb[2] = {0.0, 0.0};
  const double ybc[2] = {yb.lo, yb.hi}, susbc[2] = {susb.lo, susb.hi}, sdsbc[2] = {sdsb.lo, sdsb.hi};
#pragma unroll
  for (int J = 0; J < 2; ++J) {
    if (h == 0) {
      tb_l += ybc[J] * (1.0 - omega) / mu_c[J];
      ob_l += -ybc[J] * tau0 / mu_c[J];
    }
    if (solar) {
      const double base = omega * tau0 / mu_c[J] * Ftop;
      const double baseb = susbc[J] * pminus[J] + sdsbc[J] * pplus[J];
      pminusb[J] = susbc[J] * base;
      pplusb[J] = sdsbc[J] * base;
      if (h == 0) {
        ob_l += baseb * tau0 * Ftop / mu_c[J];
        tb_l += baseb * omega * Ftop / mu_c[J];
        fb_l += baseb * omega * tau0 / mu_c[J];
      }
    }
  }
  double taub0 = wave_sum<double>(tb_l, L) + 0.5 * g_sum;
  if (solar) taub0 -= gam_sum * gam0 / a.mu0;
  const double conb = wave_sum<double>(cb_l, L);
  double g_om = wave_sum<double>(ob_l, L) + 2.0 * kPi * conb;
  if (omega_in > 1.0) g_om = 0.0;  // clamped to 1
  const double g_F = solar ? wave_sum<double>(fb_l, L) : 0.0;
  const double g_tau = b.g_taus[task] + taub0 * ldexp(1.0, -nn);

  // Hansen normalisation reversed (src/phase_matrix.hpp:93-110)
  Mat Sp, Sn;  // Q^++ + Q^+- and Q^++ - Q^+-
  {
    double q[2] = {0.0, 0.0};
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        q[qq >> 1] += Pb_pp.v[qq][r] * Qpp.v[qq][r] + Pb_pm.v[qq][r] * Qpm.v[qq][r];
    double csumb[2];
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      const double corrb = col_reduce<double>(q[J], L);
      csumb[J] = csum[J] > 0.0 ? -corrb * corr[J] / csum[J] : 0.0;
    }
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int J = qq >> 1;
        const double wr = a.wt[bi(qq) + rw(h, r)];
        const double qbpp = csum[J] > 0.0 ? Pb_pp.v[qq][r] * corr[J] + csumb[J] * wr : Pb_pp.v[qq][r];
        const double qbpm = csum[J] > 0.0 ? Pb_pm.v[qq][r] * corr[J] + csumb[J] * wr : Pb_pm.v[qq][r];
        Sp.v[qq][r] = qbpp + qbpm;
        Sn.v[qq][r] = qbpp - qbpm;
      }
  }
  double qplusb[2] = {pplusb[0], pplusb[1]}, qminusb[2] = {pminusb[0], pminusb[1]};
  if (solar && ssol > 0.0) {  // src/phase_matrix.hpp:153-180 reversed
    const double corrb = row_sum<double>((pplusb[0] * qplus[0] + pminusb[0] * qminus[0]) +
                                         (pplusb[1] * qplus[1] + pminusb[1] * qminus[1]));
    const double sb = -corrb * corr_s / ssol;
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      qplusb[J] = pplusb[J] * corr_s + sb * wt_c[J];
      qminusb[J] = pminusb[J] * corr_s + sb * wt_c[J];
    }
  }
  const double wp[2] = {qplusb[0] + qminusb[0], qplusb[1] + qminusb[1]};
  const double wn[2] = {qplusb[0] - qminusb[0], qplusb[1] - qminusb[1]};

  // Legendre sums reversed: chi^_l = (2l+1)/(2 pi) [sum_{i,c} P_l(mu_i) P_l(mu_c)
  // (Q^++ +- Q^+-)(i, c) + sum_c P_l(mu_c) P_l(mu0) (q+^ +- q-^)(c)], + for even l
  double* chi_out = a.shared_moments ? b.g_chi + task * a.M
                                     : (b.d_mom ? b.d_mom + (gw * Lay + li) * a.M : nullptr);
  double gfr_part = 0.0;
  for (int ll = 0; ll < Lm; ++ll) {
    const bool odd = (ll & 1) != 0;
    const double* Pl = a.Pl + ll * N;
    const double plc[2] = {Pl[j], Pl[16 + j]};
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        part += Pl[bi(q) + rw(h, r)] * plc[q >> 1] * (odd ? Sn.v[q][r] : Sp.v[q][r]);
    if (solar && h == 0)
      part += a.Pl0[ll] * (plc[0] * (odd ? wn[0] : wp[0]) + plc[1] * (odd ? wn[1] : wp[1]));
    const double tot = wave_sum<double>(part, L);
    const double chisb = tot * ((2 * ll + 1) / (2.0 * kPi));
    if (mode == 3) gfr_part += chisb * (chi[ll] - 1.0) * inv1mf * inv1mf;
    if (chi_out && L.l == 0) chi_out[ll] = (mode == 3) ? chisb * inv1mf : chisb;
  }
  if (chi_out)
    for (int m = Lm + L.l; m < a.M; m += 64) chi_out[m] = 0.0;
  if (L.l == 0) {
    b.g_taus[task] = g_tau;
    b.g_omgs[task] = g_om;
    b.g_Ftop[task] = g_F;
    b.g_f[task] = (mode == 3) ? gfr_part : 0.0;
  }
}

__global__ void __launch_bounds__(256) column32_adjoint_kernel(FwdArgs a, AdjArgs b) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < a.W) w16::column_adjoint<N>(a, b, w);
}

}  // namespace w32
}  // namespace dev
}  // namespace adrt
