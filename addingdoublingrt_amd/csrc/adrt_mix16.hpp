// adrt_mix16.hpp -- mixed-precision layer kernel for N = 16 (gfx950).
//
// Same algorithm as layer16_kernel (src/doubling.hpp:29-230) with the
// doubling recursion in FP32 on v_mfma_f32_16x16x4_f32 (register rows 4h + r).
// A thin layer's transmission is I minus a small matrix (tau0 ~ 2^-16 tau), so
// FP32 T would keep only a few digits of what matters; the recursion therefore
// carries S = I - T, and every FP32 quantity is a small matrix held to full
// relative precision:
//   Q  = R^2 (I - R^2)^-1         FP32 Newton-Schulz (LU fallback), A = I - R^2,
//                                  plus one refinement step with an FP64
//                                  residual E = R R - Q + Q (R R) (FP64 matrix
//                                  cores, R and Q exact in FP64): Q += E + E Q
//   TG = T (I - R^2)^-1 = I - G,  G = S - Q + S Q
//   TGR = TG R = R - G R
//   R' = R + TGR T = R + TGR - TGR S
//   S' = I - TG T = G + S - G S
// The source vectors are carried in FP64 (their matrix products in FP32); the
// record holds R and S in FP32 (register order) and the combined sources in
// FP64; the adding kernel rebuilds T = I - S in FP64.
#pragma once

#include "adrt_fwd16.hpp"

namespace adrt {
namespace dev {
namespace w16 {

struct LdsM {
  // tiles are sized in doubles; FP32 tiles use the first half of theirs
  static constexpr int tf = 0, tf2 = TILE, td0 = 2 * TILE, pv = 3 * TILE,
                       vs = pv + N;  // 4 FP32 vector slots (2N doubles)
  static constexpr int size = vs + 2 * N;
};
// an FP32 tile (N * TS floats) fits the double tile it is given; four FP32
// vector slots of N fit 2 N doubles
static_assert(N * TS * sizeof(float) <= TILE * sizeof(double) &&
                  LdsM::tf2 + TILE <= LdsM::td0 && LdsM::td0 + TILE <= LdsM::pv &&
                  LdsM::pv + N <= LdsM::vs && LdsM::vs + 2 * N <= LdsM::size &&
                  4 * N * sizeof(float) <= 2 * N * sizeof(double),
              "mixed per-wave LDS regions overlap");

// mixed record: R, S = I - T as FP32 (register order) then s_up, s_dn as FP64
__device__ __forceinline__ float* mrec_mats(double* lay, long task) {
  return (float*)(lay + task * REC16);
}
__device__ __forceinline__ const float* mrec_mats(const double* lay, long task) {
  return (const float*)(lay + task * REC16);
}
constexpr int MREC_SRC = N * N;  // offset (doubles) of the sources in a record

// acc (FP64 layout) += X Y with X an FP64 tile and Y an FP32-layout matrix
// taken exactly into FP64 (its K index is row<float>(h, r))
__device__ __forceinline__ void mac64_f32(d4& acc, const double* X, const Mat<float>& Y,
                                          const Lane& L) {
  fence_sched();
#pragma unroll
  for (int r = 0; r < 4; ++r)
    acc = mfma(X[row<float>(L.h, r) * TS + L.j], (double)Y.v[r], acc);
  fence_sched();
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
layer16m_kernel(FwdArgs a) {
  using R = float;
  __shared__ double smem[4 * LdsM::size];
  const int wv = wave_index();
  const long task = (long)blockIdx.x * 4 + wv;
  const int L_ = a.L;
  if (task >= (long)a.W * L_) return;
  const Lane L;
  const int w = (int)(task / L_), l = (int)(task % L_);
  const long gw = (long)(a.w0 + w);
  double* sm = smem + wv * LdsM::size;
  R* tf = (R*)(sm + LdsM::tf);
  R* tf2 = (R*)(sm + LdsM::tf2);
  double* td0 = sm + LdsM::td0;
  double* pv = sm + LdsM::pv;
  R* v0 = (R*)(sm + LdsM::vs);
  R* v1 = v0 + N;
  R* v2 = v1 + N;
  R* v3 = v2 + N;
  float* om = mrec_mats(a.lay, task);
  double* os = a.lay + task * REC16 + MREC_SRC;

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
    // transparent (src/doubling.hpp:52-53) or pure absorption (:56-77), in FP64
    double trans = 1.0, strans = 0.0, su = 0.0, sd = 0.0;  // strans = 1 - trans
    if (tau > 0.0) {
      const double tex = -tau / mu_j;
      trans = (tex > -200.0) ? exp(tex) : 0.0;
      strans = (tex > -200.0) ? -expm1(tex) : 1.0;
      const double omt = 1.0 - trans;
      const double slope = mu_j * omt - 0.5 * tau * (1.0 + trans);
      su = B_bar * omt + B_d * slope;
      sd = B_bar * omt - B_d * slope;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      om[L.l * 4 + r] = 0.0f;
      om[N * N + L.l * 4 + r] = (row<R>(L.h, r) == j) ? (float)strans : 0.0f;
    }
    if (L.h == 0) {
      os[j] = su;
      os[N + j] = sd;
    }
    if (L.l == 0) a.lflag[task] = 0;
    return;
  }

  Start16<R> st;
  layer16_start<R>(a, task, w, l, gw, L, tau, omega_in, mode, td0, pv, st);
  const int nn = st.nn;
  // omega < 0.01 starts at tau0 = tau / 2^(trunc(log2 tau) + 4) (src/doubling.hpp:
  // 20-26); past the smallest mu the thin-layer T0 has negative entries that
  // the doubling amplifies by orders of magnitude and cancels, which FP32
  // cannot follow, and the adding amplifies the FP32 rounding of the layers
  // next to such a layer: all of these (marked by prep_kernel) are left to the
  // FP64 kernel (REDO pass)
  if (a.redo[task]) return;
  Mat<R> Rm, Sm;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Sm.v[r] = (float)st.Sv[r];  // S0 = I - T0 = tau0 G++, formed directly
    Rm.v[r] = (float)st.Rv[r];
  }
  double y = st.y, z = 0.0, sus = st.sus, sds = st.sds, gam = st.gam, gk = st.gk;

  for (int k = 0; k < nn; ++k) {
    Mat<R> Q, G, X;
    // ---- Q = R^2 (I - R^2)^-1 ----
    put<R>(tf, L, Rm);
    if (L.h == 0) {
      v0[j] = (float)sds;
      v1[j] = (float)sus;
    }
    wave_sync();
    mul<R>(tf, Rm, L, Q);  // R^2
    double b1 = 0.0, b2 = 0.0;
    if (a.has_solar) {
      float bs[2];
      matvecs<R, 2>(tf, v0, L, bs);
      b1 = bs[0];
      b2 = bs[1];
    }
    wave_sync();
    put<R>(tf, L, Q);
    wave_sync();
    q_form_ns<R>(Q, tf, tf2, L);  // Q0
    // FP64 residual E = R R - Q0 + Q0 (R R)
#pragma unroll
    for (int r = 0; r < 4; ++r) td0[j * TS + row<R>(L.h, r)] = (double)Rm.v[r];
    wave_sync();
    d4 P = {0.0, 0.0, 0.0, 0.0};
    mac64_f32(P, td0, Rm, L);  // R R in FP64 (FP64 layout)
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) td0[j * TS + row<R>(L.h, r)] = (double)Q.v[r];
    wave_sync();
    Mat<double> Pm{{P[0], P[1], P[2], P[3]}};
    Mat<double> QP;
    mul<double>(td0, Pm, L, QP);  // Q0 (R R)
    double E[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      E[r] = (P[r] - td0[j * TS + row<double>(L.h, r)]) + QP.v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 4; ++r) tf2[j * TS + row<double>(L.h, r)] = (float)E[r];
    wave_sync();
    {
      Mat<R> Z;
#pragma unroll
      for (int r = 0; r < 4; ++r) Z.v[r] = Q.v[r] + tf2[j * TS + row<R>(L.h, r)];
      mul_add<R>(tf2, Q, Z, L, Q);  // Q = Q0 + E + E Q0
    }
    wave_sync();
    // ---- G = S - Q + S Q ----
    put<R>(tf2, L, Sm);
    wave_sync();
    {
      Mat<R> Z;
#pragma unroll
      for (int r = 0; r < 4; ++r) Z.v[r] = Sm.v[r] - Q.v[r];
      mul_add<R>(tf2, Q, Z, L, G);
    }
    wave_sync();
    // ---- TGR = R - G R; R' = R + TGR - TGR S; S' = G + S - G S ----
    put<R>(tf2, L, G);
    const double u = z + gk * y;
    if (L.h == 0) {
      v0[j] = (float)u;
      v1[j] = (float)y;
      v2[j] = (float)(b1 + gam * sus);
      v3[j] = (float)(gam * b2 + sds);
    }
    wave_sync();
    mul<R>(tf2, Rm, L, X);  // G R
    Mat<R> TGR;
#pragma unroll
    for (int r = 0; r < 4; ++r) TGR.v[r] = Rm.v[r] - X.v[r];
    double gu, gy, gv2 = 0.0, gv3 = 0.0;
    if (a.has_solar) {
      float s4[4];
      matvecs<R, 4>(tf2, v0, L, s4);  // G u, G y, G v2, G v3
      gu = s4[0];
      gy = s4[1];
      gv2 = s4[2];
      gv3 = s4[3];
    } else {
      float s2[2];
      matvecs<R, 2>(tf2, v0, L, s2);
      gu = s2[0];
      gy = s2[1];
    }
    {
      Mat<R> Z;
#pragma unroll
      for (int r = 0; r < 4; ++r) Z.v[r] = G.v[r] + Sm.v[r];
      Mat<R> GS;
      mul<R>(tf2, Sm, L, GS);
#pragma unroll
      for (int r = 0; r < 4; ++r) Z.v[r] = Z.v[r] - GS.v[r];
      X = Z;  // S'
    }
    wave_sync();
    put<R>(tf, L, TGR);
    wave_sync();
    Mat<R> TS_;
    mul<R>(tf, Sm, L, TS_);  // TGR S
#pragma unroll
    for (int r = 0; r < 4; ++r) Rm.v[r] = (Rm.v[r] + TGR.v[r]) - TS_.v[r];
    float s2[2];
    matvecs<R, 2>(tf, v0, L, s2);  // TGR u, TGR y
    wave_sync();
    Sm = X;
    // vectors: TG x = x - G x
    const double a12 = ((u - gu) - (double)s2[0]);  // TG u - TGR u
    const double a34 = ((y - gy) + (double)s2[1]);   // TG y + TGR y
    const double znew = a12 + z - gk * y;
    const double ynew = a34 + y;
    z = znew;
    y = ynew;
    if (a.has_solar) {
      const double v2d = b1 + gam * sus, v3d = gam * b2 + sds;
      sus = (v2d - gv2) + sus;
      sds = (v3d - gv3) + gam * sds;
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // layer assembly (src/doubling.hpp:211-229), sources in FP64
  gstore<R>(om, L, Rm);
  gstore<R>(om + N * N, L, Sm);
  if (L.h == 0) {
    os[j] = (y * B_bar + z * B_d) + sus;
    os[N + j] = (y * B_bar - z * B_d) + sds;
  }
  if (L.l == 0) a.lflag[task] = 1;
}

}  // namespace w16
}  // namespace dev
}  // namespace adrt
