// adrt_mix32.hpp -- mixed-precision layer kernel for N = 32 (gfx950).
//
// The FP32 S-form doubling of the N = 16 mixed kernel (adrt_mix16.hpp) at
// N = 32: a 32 x 32 matrix is four 16 x 16 blocks per wave on
// v_mfma_f32_16x16x4_f32 (rows 4h + r of each block in lane (j, h), register
// r).  Per step, with P = R R:
//   Q  = P (I - P)^-1     FP32 Newton-Schulz in Q form (LU fallback in FP64),
//                         then one refinement with an FP64 residual
//                         E = R R - Q + Q (R R) on the FP64 matrix cores (R and
//                         Q exact in FP64): Q += E + E Q
//   G  = S - Q + S Q      (TG = T (I - P)^-1 = I - G, S = I - T)
//   TGR = R - G R,  R' = R + TGR - TGR S,  S' = G + S - G S
// with the source vectors in FP64 (their matrix products in FP32).  The
// record is FP64 -- R, and T = I - S formed in FP64 -- so the adding is the
// plain one.  Layers the prep kernel routes to FP64 (unstable doubling
// starts, their neighbours, thick strongly scattering layers) and the
// non-scattering ones are left to layer32_kernel<true>.
#pragma once

#include "adrt_wave32.hpp"

namespace adrt {
namespace dev {
namespace w32 {

typedef float f4 __attribute__((ext_vector_type(4)));

struct MatF {  // this lane's 16 entries: v[b][r] = M(16 (b & 1) + 4 h + r, 16 (b >> 1) + j)
  float v[4][4];
};
__device__ __forceinline__ int rf(int h, int r) { return 4 * h + r; }
__device__ __forceinline__ float deltaf(const Lane& L, int b, int r) {
  return (bi(b) + rf(L.h, r) == bj(b) + L.j) ? 1.0f : 0.0f;
}

// FP32 tiles: column-major, stride TS floats (N * TS floats, half a double tile)
__device__ __forceinline__ void putf(float* t, const Lane& L, const MatF& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) t[(bj(b) + L.j) * TS + bi(b) + rf(L.h, r)] = m.v[b][r];
}
__device__ __forceinline__ void getf(const float* t, const Lane& L, MatF& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m.v[b][r] = t[(bj(b) + L.j) * TS + bi(b) + rf(L.h, r)];
}
// an FP32-layout matrix into an FP64 tile (exact)
__device__ __forceinline__ void putd_f(double* t, const Lane& L, const MatF& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      t[(bj(b) + L.j) * TS + bi(b) + rf(L.h, r)] = (double)m.v[b][r];
}

// C = Z + X Y (ADD) or X Y, X in an FP32 tile, Y and Z in registers (FP32
// accumulation); two independent accumulation chains per output block column
template <bool ADD>
__device__ __forceinline__ void mulf_impl(const float* X, const MatF& Y, const MatF& Z,
                                          const Lane& L, MatF& C) {
#pragma unroll
  for (int J = 0; J < 2; ++J) {
    fence_sched();
    f4 acc[2];
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      const int b = I + 2 * J;
      acc[I] = ADD ? f4{Z.v[b][0], Z.v[b][1], Z.v[b][2], Z.v[b][3]} : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int K = 0; K < 2; ++K)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float y = Y.v[K + 2 * J][r];
#pragma unroll
        for (int I = 0; I < 2; ++I)
          acc[I] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              X[(16 * K + rf(L.h, r)) * TS + 16 * I + L.j], y, acc[I], 0, 0, 0);
      }
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) C.v[I + 2 * J][r] = acc[I][r];
    fence_sched();
  }
}
__device__ __forceinline__ void mulf(const float* X, const MatF& Y, const Lane& L, MatF& C) {
  mulf_impl<false>(X, Y, Y, L, C);
}
__device__ __forceinline__ void mulf_add(const float* X, const MatF& Y, const MatF& Z,
                                         const Lane& L, MatF& C) {
  mulf_impl<true>(X, Y, Z, L, C);
}

// C (FP64 layout) = X Y with X an FP64 tile and Y an FP32-layout matrix taken
// exactly into FP64 (K index 16 K + 4 h + r on both operands)
__device__ __forceinline__ void mul64_f32(const double* X, const MatF& Y, const Lane& L, Mat& C) {
#pragma unroll
  for (int J = 0; J < 2; ++J) {
    fence_sched();
    d4 acc[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int K = 0; K < 2; ++K)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double y = (double)Y.v[K + 2 * J][r];
#pragma unroll
        for (int I = 0; I < 2; ++I)
          acc[I] = mfma(X[(16 * K + rf(L.h, r)) * TS + 16 * I + L.j], y, acc[I]);
      }
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) C.v[I + 2 * J][r] = acc[I][r];
    fence_sched();
  }
}

// N max |E_ij| over the wave, wave-uniform
__device__ __forceinline__ float residual_sizef(const MatF& E) {
  unsigned m = 0u;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m = max(m, w16::abs_bits(E.v[b][r]));
  return w16::wave_max_abs_bits(m) * (float)N;
}

// (X v) for NV vectors in FP32 slots of 32 values (v + 32 q), X an FP32 tile:
// elements j and 16 + j, accumulated in FP32, split over the lane groups as
// matvecs
template <int NV>
__device__ __forceinline__ void matvecsf(const float* X, const float* v, const Lane& L,
                                         float (&lo)[NV], float (&hi)[NV]) {
  fence_sched();
#pragma unroll
  for (int q = 0; q < NV; ++q) lo[q] = hi[q] = 0.0f;
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const int k = 8 * L.h + kk;
    const float xl = X[k * TS + L.j], xh = X[k * TS + 16 + L.j];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const float vk = v[q * N + k];
      lo[q] = __builtin_fmaf(xl, vk, lo[q]);
      hi[q] = __builtin_fmaf(xh, vk, hi[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    lo[q] = w16::col_reduce<float>(lo[q], L);
    hi[q] = w16::col_reduce<float>(hi[q], L);
  }
  fence_sched();
}
__device__ __forceinline__ void put_vecf(float* slot, const Lane& L, const Vec& x) {
  if (L.h == 0) {
    slot[L.j] = (float)x.lo;
    slot[16 + L.j] = (float)x.hi;
  }
}

// Q = P (I - P)^-1 for P in Q on entry and in the FP32 tile tP (tX: FP32
// scratch tile), in Q form (w16::q_form_ns at N = 32): E = P - Q + P Q,
// Q <- Q + E + Q E until N max|E| <= 1e-5 N max|P|.  Where the start is poor
// (E0 = P^2 past 1/2) or after 8 steps, LU in FP64 solves Q (I - P) = P on the
// 32 x 64 system [A^T | T^T] (A = I - P, T = P) in the FP64 tiles tA and tB;
// tB holds tP and tX, so P is taken from its tile first.
__device__ __forceinline__ void q_form_ns32(MatF& Q, float* tP, float* tX, double* tA,
                                            double* tB, const Lane& L) {
  const float pmax = residual_sizef(Q);
  MatF E;
  mulf(tP, Q, L, E);  // P^2
  float e = residual_sizef(E);
  for (int it = 0; e > 1e-5f * pmax; ++it) {
    if (!(e <= 0.5f) || it == NS_MAX_STEPS) {
      MatF Pm;
      getf(tP, L, Pm);
      wave_sync();
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = bi(b) + rf(L.h, r), c = bj(b) + L.j;
          tA[i * TS + c] = (double)deltaf(L, b, r) - (double)Pm.v[b][r];  // A^T column-major
          tB[i * TS + c] = (double)Pm.v[b][r];                             // T^T column-major
        }
      wave_sync();
      gauss_jordan(tA, tB, L);
      wave_sync();
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Q.v[b][r] = (float)tB[(bi(b) + rf(L.h, r)) * TS + bj(b) + L.j];  // X = (X^T)^T
      wave_sync();
      return;
    }
    putf(tX, L, Q);
    wave_sync();
    {
      MatF Z;
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) Z.v[b][r] = Q.v[b][r] + E.v[b][r];
      mulf_add(tX, E, Z, L, Q);  // Q + E + Q E
    }
    wave_sync();
    MatF PQ, Pm;
    mulf(tP, Q, L, PQ);
    getf(tP, L, Pm);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) E.v[b][r] = (Pm.v[b][r] - Q.v[b][r]) + PQ.v[b][r];
    e = residual_sizef(E);
  }
  wave_sync();
}

// LDS per wave (doubles): an FP64 tile, an FP64 tile holding two FP32 tiles,
// eight FP32 vector slots of 32 (in 4 N doubles) -- the same extent as SLOT
struct LdsM32 {
  static constexpr int td = 0, tb = TILE, fv = 2 * TILE, size = 2 * TILE + 4 * N;
};
static_assert(2 * N * TS * sizeof(float) <= TILE * sizeof(double) &&
                  LdsM32::td + TILE <= LdsM32::tb && LdsM32::tb + TILE <= LdsM32::fv &&
                  8 * N * sizeof(float) <= 4 * N * sizeof(double) && LdsM32::size <= SLOT,
              "N = 32 mixed per-wave LDS regions overlap");

__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2, 8)))
layer32m_kernel(FwdArgs a) {
  __shared__ double smem[WAVES * SLOT];
  const int wv = wave_index();
  const long task = (long)blockIdx.x * WAVES + wv;
  const int Lay = a.L;
  if (task >= (long)a.W * Lay) return;
  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  // routed and non-scattering layers: layer32_kernel<true>
  if (a.redo[task] || !(tau > 0.0 && omega_in > 0.0)) return;
  const Lane L;
  const int j = L.j, h = L.h;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  double* sm = smem + wv * SLOT;
  double* td = sm + LdsM32::td;
  double* tb = sm + LdsM32::tb;
  float* tf = (float*)tb;
  float* tf2 = tf + N * TS;
  float* fv = (float*)(sm + LdsM32::fv);
  const int mode = a.mode[task];
  const double Btop = a.B ? a.B[gw * (Lay + 1) + lev(a, l)] : 0.0;
  const double Bbot = a.B ? a.B[gw * (Lay + 1) + lev(a, l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (Bbot - Btop) / tau;

  MatF Rm, Sm;
  Vec y, z{0.0, 0.0}, sus, sds;
  double gam, gk;
  int nn;
  {
    Start32 st;
    start32<float, true>(a, task, w, l, gw, L, tau, omega_in, mode, st);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Rm.v[b][r] = (float)st.R.v[b][r];
        Sm.v[b][r] = (float)st.T.v[b][r];  // S0 = tau0 G++, formed directly
      }
    y = st.y;
    sus = st.sus;
    sds = st.sds;
    gam = st.gam;
    gk = st.gk;
    nn = st.nn;
  }
  const bool solar = a.has_solar != 0;

  for (int k = 0; k < nn; ++k) {
    MatF Q, G, X;
    // ---- Q = R^2 (I - R^2)^-1 ----
    putf(tf, L, Rm);
    put_vecf(fv, L, sds);
    put_vecf(fv + N, L, sus);
    wave_sync();
    mulf(tf, Rm, L, Q);  // R^2
    Vec b1{0.0, 0.0}, b2{0.0, 0.0};
    if (solar) {
      float lo[2], hi[2];
      matvecsf<2>(tf, fv, L, lo, hi);  // R s_dn, R s_up
      b1 = Vec{lo[0], hi[0]};
      b2 = Vec{lo[1], hi[1]};
    }
    wave_sync();
    putf(tf, L, Q);
    wave_sync();
    q_form_ns32(Q, tf, tf2, td, tb, L);  // Q0
    // FP64 residual E = R R - Q0 + Q0 (R R), then Q = Q0 + E + E Q0
    putd_f(td, L, Rm);
    wave_sync();
    Mat P;
    mul64_f32(td, Rm, L, P);  // R R (FP64 layout)
    wave_sync();
    putd_f(td, L, Q);
    wave_sync();
    Mat QP;
    mul(td, P, L, QP);  // Q0 (R R)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = (bj(b) + j) * TS + bi(b) + rw(h, r);  // FP64-layout position
        tf2[o] = (float)((P.v[b][r] - td[o]) + QP.v[b][r]);
      }
    wave_sync();
    {
      MatF Z;
      getf(tf2, L, Z);  // E in the FP32 layout
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) Z.v[b][r] = Q.v[b][r] + Z.v[b][r];
      mulf_add(tf2, Q, Z, L, Q);
    }
    wave_sync();
    // ---- G = S - Q + S Q ----
    putf(tf2, L, Sm);
    wave_sync();
    {
      MatF Z;
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) Z.v[b][r] = Sm.v[b][r] - Q.v[b][r];
      mulf_add(tf2, Q, Z, L, G);
    }
    wave_sync();
    // ---- TGR = R - G R; R' = R + TGR - TGR S; S' = G + S - G S ----
    putf(tf2, L, G);
    const Vec u{z.lo + gk * y.lo, z.hi + gk * y.hi};
    const Vec v2d{b1.lo + gam * sus.lo, b1.hi + gam * sus.hi};
    const Vec v3d{gam * b2.lo + sds.lo, gam * b2.hi + sds.hi};
    put_vecf(fv, L, u);
    put_vecf(fv + N, L, y);
    put_vecf(fv + 2 * N, L, v2d);
    put_vecf(fv + 3 * N, L, v3d);
    wave_sync();
    mulf(tf2, Rm, L, X);  // G R
    MatF TGR;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) TGR.v[b][r] = Rm.v[b][r] - X.v[b][r];
    float glo[4] = {0.f, 0.f, 0.f, 0.f}, ghi[4] = {0.f, 0.f, 0.f, 0.f};
    if (solar) {
      matvecsf<4>(tf2, fv, L, glo, ghi);  // G u, G y, G v2, G v3
    } else {
      float lo2[2], hi2[2];
      matvecsf<2>(tf2, fv, L, lo2, hi2);
      glo[0] = lo2[0];
      glo[1] = lo2[1];
      ghi[0] = hi2[0];
      ghi[1] = hi2[1];
    }
    {
      MatF GS;
      mulf(tf2, Sm, L, GS);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) X.v[b][r] = (G.v[b][r] + Sm.v[b][r]) - GS.v[b][r];  // S'
    }
    wave_sync();
    putf(tf, L, TGR);
    wave_sync();
    {
      MatF TS_;
      mulf(tf, Sm, L, TS_);  // TGR S
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) Rm.v[b][r] = (Rm.v[b][r] + TGR.v[b][r]) - TS_.v[b][r];
    }
    float tlo[2], thi[2];
    matvecsf<2>(tf, fv, L, tlo, thi);  // TGR u, TGR y
    wave_sync();
    Sm = X;
    // vectors: TG x = x - G x
    const Vec a12{(u.lo - glo[0]) - (double)tlo[0], (u.hi - ghi[0]) - (double)thi[0]};
    const Vec a34{(y.lo - glo[1]) + (double)tlo[1], (y.hi - ghi[1]) + (double)thi[1]};
    const Vec znew{a12.lo + z.lo - gk * y.lo, a12.hi + z.hi - gk * y.hi};
    const Vec ynew{a34.lo + y.lo, a34.hi + y.hi};
    z = znew;
    y = ynew;
    if (solar) {
      sus = Vec{(v2d.lo - glo[2]) + sus.lo, (v2d.hi - ghi[2]) + sus.hi};
      sds = Vec{(v3d.lo - glo[3]) + gam * sds.lo, (v3d.hi - ghi[3]) + gam * sds.hi};
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // layer assembly (src/doubling.hpp:211-229): FP64 record, T = I - S in FP64
  constexpr int REC = 2 * N * N + 2 * N;
  double* out = a.lay + task * REC;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = (bj(b) + j) * N + bi(b) + rf(h, r);  // column-major record
      out[o] = (double)Rm.v[b][r];
      out[N * N + o] = (double)deltaf(L, b, r) - (double)Sm.v[b][r];
    }
  if (h == 0) {
    out[2 * N * N + j] = (y.lo * B_bar + z.lo * B_d) + sus.lo;
    out[2 * N * N + 16 + j] = (y.hi * B_bar + z.hi * B_d) + sus.hi;
    out[2 * N * N + N + j] = (y.lo * B_bar - z.lo * B_d) + sds.lo;
    out[2 * N * N + N + 16 + j] = (y.hi * B_bar - z.hi * B_d) + sds.hi;
    if (a.basis) {
      double* bs = a.basis + task * 2 * N;
      bs[j] = y.lo;
      bs[16 + j] = y.hi;
      bs[N + j] = z.lo;
      bs[N + 16 + j] = z.hi;
    }
  }
  if (j == 0 && h == 0) a.lflag[task] = 1;
}

}  // namespace w32
}  // namespace dev
}  // namespace adrt
