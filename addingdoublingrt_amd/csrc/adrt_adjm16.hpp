// adrt_adjm16.hpp -- reverse layer pass of the mixed-precision form (N = 16).
//
// Differentiates the FP32 S-form doubling of layer16m_kernel (adrt_mix16.hpp):
// per step, with P = R R, A = I - P, Q = P A^-1 (so A^-1 = I + Q),
//   G = S - Q + S Q,  TGR = R - G R,  R' = R + TGR - TGR S,  S' = G + S - G S,
//   u = z + g y,  z' = (u - G u) - TGR u + z - g y,  y' = (y - G y) + TGR y + y,
//   s_up' = (v2 - G v2) + s_up,  s_dn' = (v3 - G v3) + gamma s_dn,
//   v2 = R s_dn + gamma s_up,  v3 = gamma R s_up + s_dn.
// Reversed, with X^ the cotangent of X:
//   TGR^ = R'^ - R'^ S^T - z'^ u^T + y'^ y^T
//   G^   = S'^ - S'^ S^T - TGR^ R^T - z'^ u^T - y'^ y^T - s_up'^ v2^T - s_dn'^ v3^T
//   S^   = S'^ - G^T S'^ - TGR^T R'^ + G^ + G^ Q^T
//   Q^   = -G^ + S^T G^,   P^ = (I + Q)^T Q^ (I + Q)^T
//   R^   = R'^ + TGR^ - G^T TGR^ + P^ R^T + R^T P^ + v2^ s_dn^T + gamma v3^ s_up^T
// (v2^ = s_up'^ - G^T s_up'^, v3^ = s_dn'^ - G^T s_dn'^): matrix products only,
// on the FP32 matrix cores, from per-step checkpoints of R, S and Q (FP32) and
// the FP64 vectors.  The doubling is recomputed as the forward does it, but
// with Q's refinement step in FP32 (the forward takes its residual in FP64).
// Layers the forward left to the FP64 kernel (redo flag) are skipped here and
// handled by layer16_adjoint_kernel<true>.
#pragma once

#include "adrt_adj16.hpp"
#include "adrt_mix16.hpp"

namespace adrt {
namespace dev {
namespace w16 {

struct LdsAM {
  // in doubles: eight FP32 tiles (half a double tile each), the first two of
  // which hold the FP64 tile of the thin-layer start before the doubling loops
  // use them; eight FP32 vector slots, one FP64 column-sum slot (9.4 KB per
  // wave, so four 4-wave workgroups fit a CU's 160 KB)
  static constexpr int TF = TILE / 2 + 1;
  static constexpr int td = 0, tf0 = 0, vs = 8 * TF, csl = vs + 4 * N;
  static constexpr int size = csl + N;
};
static_assert(N * TS * sizeof(float) <= LdsAM::TF * sizeof(double) &&
                  LdsAM::td + TILE <= LdsAM::tf0 + 2 * LdsAM::TF &&
                  LdsAM::tf0 + 8 * LdsAM::TF <= LdsAM::vs &&
                  8 * N * sizeof(float) <= 4 * N * sizeof(double) &&
                  LdsAM::vs + 4 * N <= LdsAM::csl && LdsAM::csl + N <= LdsAM::size,
              "mixed reverse per-wave LDS regions overlap");

// per-step checkpoint (within the SLAB stride): R, S, Q in FP32 register
// order, then y, z, s_up, s_dn (FP64, element j) and gamma
constexpr int MS_Y = 3 * N * N / 2, MS_Z = MS_Y + N, MS_SU = MS_Z + N, MS_SD = MS_SU + N,
              MS_GAM = MS_SD + N;
static_assert(MS_GAM < SLAB, "mixed checkpoint exceeds the slab stride");

__device__ void layer16m_adjoint_task(const FwdArgs& a, const AdjArgs& b, long task, const Lane& L,
                                      double* sm, double* slab) {
  using R = float;
  const int Lay = a.L;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  const int j = L.j;
  double* td = sm + LdsAM::td;
  R* tf[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) tf[q] = (R*)(sm + LdsAM::tf0 + q * LdsAM::TF);
  R* fs = (R*)(sm + LdsAM::vs);  // 8 FP32 slots of N
  double* csl = sm + LdsAM::csl;
  R* tS = tf[0];
  R* tR = tf[1];
  R* tQ = tf[2];
  R* tG = tf[3];
  R* tT = tf[4];
  R* tA = tf[5];
  R* tB = tf[6];
  R* tC = tf[7];

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
  auto finish = [&](int nchi) {
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

  // the thin-layer start as the forward computed it; its reverse is
  // start16_adjoint_kernel<true>'s (in the FP32 register order)
  Mat<R> Rm, Sm;
  double y, z = 0.0, sus, sds, gam, gk, tau0;
  int nn;
  {
    Start16<R> st0;
    layer16_start<R>(a, task, w, l, gw, L, tau, omega_in, mode, td, csl, st0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Sm.v[r] = (float)st0.Sv[r];  // S0 = tau0 G++, formed directly
      Rm.v[r] = (float)st0.Rv[r];
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

  // ---- the forward S-form doubling, recomputed with checkpoints ----
  for (int k = 0; k < nn; ++k) {
    double* st = slab + (long)k * SLAB;
    float* sf = (float*)st;
    gstore<R>(sf, L, Rm);
    gstore<R>(sf + N * N, L, Sm);
    if (L.h == 0) {
      st[MS_Y + j] = y;
      st[MS_Z + j] = z;
      st[MS_SU + j] = sus;
      st[MS_SD + j] = sds;
    }
    if (L.l == 0) st[MS_GAM] = gam;
    Mat<R> Q, G, X;
    put<R>(tA, L, Rm);
    if (L.h == 0) {
      fs[j] = (float)sds;
      fs[N + j] = (float)sus;
    }
    wave_sync();
    mul<R>(tA, Rm, L, Q);  // R^2
    double b1 = 0.0, b2 = 0.0;
    if (solar) {
      float bs[2];
      matvecs<R, 2>(tA, fs, L, bs);
      b1 = bs[0];
      b2 = bs[1];
    }
    wave_sync();
    put<R>(tA, L, Q);
    wave_sync();
    q_form_ns<R>(Q, tA, tB, L);
    // one refinement step in FP32, E = P - Q + Q P, Q <- Q + E + E Q: Q to
    // FP32 rounding of the P = R R of this FP32 recursion (the forward's
    // refinement takes E in FP64 for its 1e-6 flux accuracy; the gradients'
    // 1e-4 needs only FP32 consistency)
    {
      Mat<R> Pm, QP, Zm;
      get<R>(tA, L, Pm);
      put<R>(tB, L, Q);
      wave_sync();
      mul<R>(tB, Pm, L, QP);
#pragma unroll
      for (int r = 0; r < 4; ++r) Pm.v[r] = (Pm.v[r] - Q.v[r]) + QP.v[r];  // E
      wave_sync();
      put<R>(tB, L, Pm);
      wave_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) Zm.v[r] = Q.v[r] + Pm.v[r];
      mul_add<R>(tB, Q, Zm, L, Q);
    }
    gstore<R>(sf + 2 * N * N, L, Q);
    wave_sync();
    put<R>(tB, L, Sm);
    wave_sync();
    {
      Mat<R> Zm;
#pragma unroll
      for (int r = 0; r < 4; ++r) Zm.v[r] = Sm.v[r] - Q.v[r];
      mul_add<R>(tB, Q, Zm, L, G);
    }
    wave_sync();
    put<R>(tB, L, G);
    const double u = z + gk * y;
    if (L.h == 0) {
      fs[j] = (float)u;
      fs[N + j] = (float)y;
      fs[2 * N + j] = (float)(b1 + gam * sus);
      fs[3 * N + j] = (float)(gam * b2 + sds);
    }
    wave_sync();
    mul<R>(tB, Rm, L, X);
    Mat<R> TGR;
#pragma unroll
    for (int r = 0; r < 4; ++r) TGR.v[r] = Rm.v[r] - X.v[r];
    double gu, gy, gv2 = 0.0, gv3 = 0.0;
    if (solar) {
      float s4[4];
      matvecs<R, 4>(tB, fs, L, s4);
      gu = s4[0];
      gy = s4[1];
      gv2 = s4[2];
      gv3 = s4[3];
    } else {
      float s2[2];
      matvecs<R, 2>(tB, fs, L, s2);
      gu = s2[0];
      gy = s2[1];
    }
    const bool last = k == nn - 1;  // the last R', S' are never read back
    if (!last) {
      Mat<R> GS;
      mul<R>(tB, Sm, L, GS);
#pragma unroll
      for (int r = 0; r < 4; ++r) X.v[r] = (G.v[r] + Sm.v[r]) - GS.v[r];  // S'
    }
    wave_sync();
    put<R>(tA, L, TGR);
    wave_sync();
    if (!last) {
      Mat<R> TS_;
      mul<R>(tA, Sm, L, TS_);
#pragma unroll
      for (int r = 0; r < 4; ++r) Rm.v[r] = (Rm.v[r] + TGR.v[r]) - TS_.v[r];
    }
    float s2[2];
    matvecs<R, 2>(tA, fs, L, s2);
    wave_sync();
    if (!last) Sm = X;
    const double znew = ((u - gu) - (double)s2[0]) + z - gk * y;
    const double ynew = ((y - gy) + (double)s2[1]) + y;
    z = znew;
    y = ynew;
    if (solar) {
      const double v2d = b1 + gam * sus, v3d = gam * b2 + sds;
      sus = (v2d - gv2) + sus;
      sds = (v3d - gv3) + gam * sds;
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // ---- layer assembly reversed (sources are FP64) ----
  const double sub = c.sub, sdb = c.sdb;
  double yb = (sub + sdb) * B_bar, zb = (sub - sdb) * B_d, susb = sub, sdsb = sdb;
  {
    const double gbar = colsum<double>((sub + sdb) * y, L, csl);
    const double gd = colsum<double>((sub - sdb) * z, L, csl);
    g_Bt = 0.5 * gbar - gd / tau;
    g_Bb = 0.5 * gbar + gd / tau;
    g_tau = -gd * B_d / tau;
  }
  double gamb = 0.0, gb = 0.0;
  Mat<R> Rbar, Sbar;  // cotangents of R and S = I - T
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = row<R>(L.h, r);
    Rbar.v[r] = (float)cot_R(c, i);
    Sbar.v[r] = (float)(-cot_T(c, i));
  }

  // ---- S-form doubling steps in reverse ----
  for (int k = nn - 1; k >= 0; --k) {
    const double* st = slab + (long)k * SLAB;
    const float* sf = (const float*)st;
    Mat<R> Rk, Sk, Qk, G, TGR, St, Rt, Qt, X;
    gload<R>(sf, L, Rk);
    gload<R>(sf + N * N, L, Sk);
    gload<R>(sf + 2 * N * N, L, Qk);
    const double yk = st[MS_Y + j], zk = st[MS_Z + j], susk = st[MS_SU + j], sdsk = st[MS_SD + j];
    const double gm = st[MS_GAM];
    const double g = 0.5 * tau0 * ldexp(1.0, k);
    put<R>(tS, L, Sk);
    put<R>(tR, L, Rk);
    put<R>(tQ, L, Qk);
    if (L.h == 0) {
      fs[j] = (float)sdsk;
      fs[N + j] = (float)susk;
      fs[2 * N + j] = (float)zb;
      fs[3 * N + j] = (float)yb;
      fs[4 * N + j] = (float)susb;
      fs[5 * N + j] = (float)sdsb;
    }
    wave_sync();
    {
      Mat<R> Zm;
#pragma unroll
      for (int r = 0; r < 4; ++r) Zm.v[r] = Sk.v[r] - Qk.v[r];
      mul_add<R>(tS, Qk, Zm, L, G);  // G = S - Q + S Q
    }
    double b2 = 0.0;
    double v2 = 0.0, v3 = 0.0;
    if (solar) {
      float bs[2];
      matvecs<R, 2>(tR, fs, L, bs);  // R s_dn, R s_up
      v2 = (double)bs[0] + gm * susk;
      v3 = gm * (double)bs[1] + sdsk;
      b2 = bs[1];
    }
    get_t<R>(tS, L, St);
    get_t<R>(tR, L, Rt);
    get_t<R>(tQ, L, Qt);
    put<R>(tG, L, G);
    wave_sync();
    mul<R>(tG, Rk, L, X);
#pragma unroll
    for (int r = 0; r < 4; ++r) TGR.v[r] = Rk.v[r] - X.v[r];
    put<R>(tT, L, TGR);
    const double u = zk + g * yk;
    // vector recursions reversed
    const double tGz = tmatvec<R>(G, fs + 2 * N, L), tGy = tmatvec<R>(G, fs + 3 * N, L);
    const double tRz = tmatvec<R>(TGR, fs + 2 * N, L), tRy = tmatvec<R>(TGR, fs + 3 * N, L);
    const double ub = (zb - tGz) - tRz;
    const double yb_n = ((2.0 * yb - g * zb) - tGy + tRy) + g * ub;
    const double zb_n = zb + ub;
    gb = 2.0 * gb + (ub - zb) * yk;
    double v2b = 0.0, v3b = 0.0, b1b = 0.0, b2b = 0.0, susb_n = susb, sdsb_n = sdsb;
    if (solar) {
      v2b = susb - tmatvec<R>(G, fs + 4 * N, L);
      v3b = sdsb - tmatvec<R>(G, fs + 5 * N, L);
      b1b = v2b;
      b2b = gm * v3b;
      wave_sync();
      if (L.h == 0) {
        fs[6 * N + j] = (float)b1b;
        fs[7 * N + j] = (float)b2b;
      }
      wave_sync();
      susb_n = (susb + gm * v2b) + tmatvec<R>(Rk, fs + 7 * N, L);
      sdsb_n = (gm * sdsb + v3b) + tmatvec<R>(Rk, fs + 6 * N, L);
      gamb = 2.0 * gm * gamb + ((sdsb * sdsk + v2b * susk) + v3b * b2);
    }
    // matrices reversed
    put<R>(tA, L, Rbar);
    put<R>(tB, L, Sbar);
    wave_sync();
    f4 acc = zero4<R>();
    mac<R>(acc, tA, St, L);  // R'^ S^T
    Mat<R> TGRb;
    {
      f4 o = zero4<R>();
      outer4<R>(o, (float)-zb, (float)yb, 0.0f, 0.0f, (float)u, (float)yk, 0.0f, 0.0f, L);
#pragma unroll
      for (int r = 0; r < 4; ++r) TGRb.v[r] = (Rbar.v[r] - acc[r]) + o[r];
    }
    put<R>(tC, L, TGRb);
    wave_sync();
    Mat<R> Gb;
    {
      f4 m1 = zero4<R>(), m2 = zero4<R>(), o = zero4<R>();
      mac<R>(m1, tB, St, L);  // S'^ S^T
      mac<R>(m2, tC, Rt, L);  // TGR^ R^T
      outer4<R>(o, (float)-zb, (float)-yb, (float)-susb, (float)-sdsb, (float)u, (float)yk,
                (float)v2, (float)v3, L);
#pragma unroll
      for (int r = 0; r < 4; ++r) Gb.v[r] = ((Sbar.v[r] - m1[r]) - m2[r]) + o[r];
    }
    wave_sync();
    put<R>(tA, L, Gb);  // R'^ no longer needed as a tile
    wave_sync();
    Mat<R> Sbar_n, Qb;
    {
      f4 m1 = zero4<R>(), m2 = zero4<R>(), m3 = zero4<R>(), m4 = zero4<R>();
      mac_t<R>(m1, tT, Rbar, L);  // TGR^T R'^
      mac_t<R>(m2, tG, Sbar, L);  // G^T S'^
      mac<R>(m3, tA, Qt, L);      // G^ Q^T
      mac_t<R>(m4, tS, Gb, L);    // S^T G^
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Sbar_n.v[r] = (((Sbar.v[r] - m2[r]) - m1[r]) + Gb.v[r]) + m3[r];
        Qb.v[r] = m4[r] - Gb.v[r];
      }
    }
    wave_sync();
    put<R>(tB, L, Qb);
    wave_sync();
    Mat<R> Zq, Pb;
    {
      f4 m = zero4<R>();
      mac<R>(m, tB, Qt, L);  // Q^ Q^T
#pragma unroll
      for (int r = 0; r < 4; ++r) Zq.v[r] = Qb.v[r] + m[r];
      f4 m2 = acc_of<R>(Zq);
      mac_t<R>(m2, tQ, Zq, L);  // Z + Q^T Z
      Pb = mat_of<R>(m2);
    }
    wave_sync();
    put<R>(tB, L, Pb);
    wave_sync();
    {
      f4 m1 = zero4<R>(), m2 = zero4<R>(), m3 = zero4<R>(), o = zero4<R>();
      mac_t<R>(m1, tG, TGRb, L);  // G^T TGR^
      mac<R>(m2, tB, Rt, L);      // P^ R^T
      mac_t<R>(m3, tR, Pb, L);    // R^T P^
      outer4<R>(o, (float)b1b, (float)b2b, 0.0f, 0.0f, (float)sdsk, (float)susk, 0.0f, 0.0f, L);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Rbar.v[r] = ((((Rbar.v[r] + TGRb.v[r]) - m1[r]) + m2[r]) + m3[r]) + o[r];
    }
    Sbar = Sbar_n;
    yb = yb_n;
    zb = zb_n;
    susb = susb_n;
    sdsb = sdsb_n;
    wave_sync();
  }

  // ---- cotangents of the doubling start, for start16_adjoint_kernel ----
  const double g_sum = colsum<double>(gb, L, csl);
  const double gam_sum = solar ? colsum<double>(gamb, L, csl) : 0.0;
  double* sc = b.scot + task * SCOT;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    sc[L.l * 4 + r] = (double)Rbar.v[r];
    sc[N * N + L.l * 4 + r] = -(double)Sbar.v[r];  // T0 = I - S0
  }
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

constexpr int MREV_WAVES_PER_SIMD = 4;  // register budget 128 per lane

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MREV_WAVES_PER_SIMD, 8)))
layer16m_adjoint_kernel(FwdArgs a, AdjArgs b) {
  __shared__ double smem[4 * LdsAM::size];
  const int wv = wave_index();
  const int gwave = blockIdx.x * 4 + wv;
  const Lane L;
  double* sm = smem + wv * LdsAM::size;
  double* slab = b.slab + (long)gwave * slab_stride(b.nn_cap);
  const long ntask = (long)a.W * a.L;
  for_tasks<TASK_CHUNK>(b.queue, ntask, [&](long task) {
    if (a.redo[task]) return;  // FP64 layer (layer16_adjoint_kernel<true>)
    layer16m_adjoint_task(a, b, task, L, sm, slab);
  });
}

}  // namespace w16
}  // namespace dev
}  // namespace adrt
