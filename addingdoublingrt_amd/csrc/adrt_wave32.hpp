// N = 32 on the matrix cores: the doubling with one wave per (wavenumber,
// layer), the adding sweeps with one wave per wavenumber.
//
// A 32 x 32 matrix is four 16 x 16 blocks in the register layout of the N = 16
// path (adrt_wave16.hpp): block b = I + 2J holds rows 16I.., columns 16J..,
// and lane l = j + 16h holds entries (16I + h + 4r, 16J + j), r = 0..3, of
// every block.  A product C = X Y reads X from a padded (stride 33)
// column-major LDS tile and Y from registers: each output block is eight
// v_mfma_f64_16x16x4 (K = 32), 32 per product.  Vectors of 32 are held as the
// pair (element j, element 16 + j) in every lane of column j.
//
// The layer kernel restates the doubling of src/doubling.hpp:29-230 and the
// phase matrices of src/phase_matrix.hpp:61-181 as the column-slice layer
// kernel does (adrt_fwd.hpp) and writes the same records; the adding kernel
// restates src/adding.hpp:69-258 and src/solver.cpp:516-700 as the
// column-slice adding kernel does.  Every (I - P)^-1 is a Newton-Schulz
// inverse from the Neumann start; where that start is poor, LU with
// partial pivoting solves against the right-hand side in LDS, one augmented
// column per lane.
#pragma once

#include "adrt_fwd16.hpp"
#include "adrt_wave16.hpp"

namespace adrt {
namespace dev {
namespace w32 {

using w16::d4;
using w16::fence_sched;
using w16::Lane;
using w16::mfma;
using w16::NS_CUBIC;
using w16::NS_DONE;
using w16::NS_MAX_STEPS;
using w16::NS_NEWTON;
using w16::readlane;
using w16::row_max;
using w16::row_sum;
using w16::shfl;
using w16::wave_index;

constexpr int N = 32;
constexpr int TS = 33;          // LDS tile column stride
constexpr int TILE = N * TS;    // doubles per tile
constexpr int WAVES = 2;        // waves per 128-thread block
constexpr int SLOT = 2 * TILE + 4 * N;  // LDS doubles per wave: two tiles, four vector slots
// every tile access indexes tile[k * TS + i], 0 <= i, k < N; the LU fallback
// keeps its 32 x 64 augmented system in the wave's two tiles
static_assert((N - 1) * TS + (N - 1) < TILE && 2 * TILE + 4 * N <= SLOT,
              "N = 32 per-wave LDS regions overlap");

struct Mat {  // this lane's 16 entries: v[b][r] = M(16 (b & 1) + h + 4r, 16 (b >> 1) + j)
  double v[4][4];
};
struct Vec {  // elements j and 16 + j
  double lo, hi;
};

__device__ __forceinline__ int rw(int h, int r) { return h + 4 * r; }
__device__ __forceinline__ int bi(int b) { return 16 * (b & 1); }
__device__ __forceinline__ int bj(int b) { return 16 * (b >> 1); }
__device__ __forceinline__ double delta(const Lane& L, int b, int r) {
  return (bi(b) + rw(L.h, r) == bj(b) + L.j) ? 1.0 : 0.0;
}

__device__ __forceinline__ void put(double* t, const Lane& L, const Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) t[(bj(b) + L.j) * TS + bi(b) + rw(L.h, r)] = m.v[b][r];
}
__device__ __forceinline__ void get(const double* t, const Lane& L, Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m.v[b][r] = t[(bj(b) + L.j) * TS + bi(b) + rw(L.h, r)];
}
// the transpose of m into the tile (column-major M^T)
__device__ __forceinline__ void put_t(double* t, const Lane& L, const Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) t[(bi(b) + rw(L.h, r)) * TS + bj(b) + L.j] = m.v[b][r];
}
// the transpose of the tile's matrix
__device__ __forceinline__ void get_t(const double* t, const Lane& L, Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m.v[b][r] = t[(bi(b) + rw(L.h, r)) * TS + bj(b) + L.j];
}

// C = Z + X Y (ADD) or X Y, X in a tile, Y and Z in registers; C may alias Y or Z
template <bool ADD>
__device__ __forceinline__ void mul_impl(const double* X, const Mat& Y, const Mat& Z,
                                         const Lane& L, Mat& C) {
#pragma unroll
  for (int J = 0; J < 2; ++J) {
    fence_sched();
    d4 acc[2];
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      const int b = I + 2 * J;
      acc[I] = ADD ? d4{Z.v[b][0], Z.v[b][1], Z.v[b][2], Z.v[b][3]} : d4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int K = 0; K < 2; ++K)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double y = Y.v[K + 2 * J][r];
#pragma unroll
        for (int I = 0; I < 2; ++I)
          acc[I] = mfma(X[(16 * K + rw(L.h, r)) * TS + 16 * I + L.j], y, acc[I]);
      }
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) C.v[I + 2 * J][r] = acc[I][r];
    fence_sched();
  }
}
__device__ __forceinline__ void mul(const double* X, const Mat& Y, const Lane& L, Mat& C) {
  mul_impl<false>(X, Y, Y, L, C);
}
__device__ __forceinline__ void mul_add(const double* X, const Mat& Y, const Mat& Z,
                                        const Lane& L, Mat& C) {
  mul_impl<true>(X, Y, Z, L, C);
}

// (X v) for NV vectors stored in LDS slots of 32 values, v + 32 q: lane group
// h takes k = 8h..8h+7, the four partial sums are added across the column
template <int NV>
__device__ __forceinline__ void matvecs(const double* X, const double* v, const Lane& L,
                                        Vec (&s)[NV]) {
  fence_sched();
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = Vec{0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const int k = 8 * L.h + kk;
    const double xl = X[k * TS + L.j], xh = X[k * TS + 16 + L.j];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double vk = v[q * N + k];
      s[q].lo = fma_(xl, vk, s[q].lo);
      s[q].hi = fma_(xh, vk, s[q].hi);
    }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q)
    s[q] = Vec{w16::col_reduce<double>(s[q].lo, L), w16::col_reduce<double>(s[q].hi, L)};
  fence_sched();
}

__device__ __forceinline__ void put_vec(double* slot, const Lane& L, const Vec& x) {
  if (L.h == 0) {
    slot[L.j] = x.lo;
    slot[16 + L.j] = x.hi;
  }
}

// N max |E_ij| over the wave, wave-uniform
__device__ __forceinline__ float residual_size(const Mat& E) {
  unsigned m = 0u;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m = max(m, w16::abs_bits((float)E.v[b][r]));
  return w16::wave_max_abs_bits(m) * (float)N;
}

// X = (I - P)^-1 by Newton-Schulz for P in registers and in the tile tP (tX:
// scratch tile), the iteration of the N = 16 path (inverse_ns): X0 = I + P,
// E0 = P^2, X <- X + X E (or X + X E + X E E once N max|E| <= 1e-5),
// E = I - X + P X.  Returns true, X unset, where the start is poor
// (N max|E0| > 1/2) or the iteration has not finished in 8 steps: the caller
// then solves with its right-hand side.
__device__ __forceinline__ bool inverse_ns(const Mat& P, const double* tP, double* tX, Mat& X,
                                           const Lane& L) {
  Mat E;
  mul(tP, P, L, E);
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) X.v[b][r] = delta(L, b, r) + P.v[b][r];
  float e = residual_size(E);
  for (int it = 0; e > NS_DONE; ++it) {
    if (!(e <= 0.5f) || it == NS_MAX_STEPS) return true;
    put(tX, L, X);
    wave_sync();
    if (e <= NS_NEWTON) {
      mul_add(tX, E, X, L, X);
      wave_sync();
      break;
    }
    Mat XE;
    mul(tX, E, L, XE);
    wave_sync();
    if (e <= NS_CUBIC) {
      put(tX, L, XE);
      wave_sync();
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) X.v[b][r] += XE.v[b][r];
      mul_add(tX, E, X, L, X);
      wave_sync();
      break;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) X.v[b][r] += XE.v[b][r];
    Mat PX;
    mul(tP, X, L, PX);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) E.v[b][r] = (delta(L, b, r) - X.v[b][r]) + PX.v[b][r];
    e = residual_size(E);
  }
  wave_sync();
  return false;
}

// LU with partial pivoting on the 32 x 64 augmented system [A^T | T^T], A^T
// column-major in tA and T^T in tB; lane c owns augmented column c.  The pivot
// of step k is the first largest |A^T(i, k)|, i >= k (the pivots of the
// reference's LU of A^T, src/matrix.hpp:120-124); the pivot row is scaled to a
// unit diagonal and eliminated from the rows below, then back substitution
// clears the rows above (eliminating above during the forward sweep,
// Gauss-Jordan, leaves a residual larger by up to cond(A)).  Every read of a
// column that its owner rewrites in the same pass is separated from the write
// by a wave barrier, chunk by chunk.  On return tB holds X^T with X A = T.
__device__ __forceinline__ void gauss_jordan(double* tA, double* tB, const Lane& L) {
  double* col = (L.l < 32 ? tA : tB) + (L.l & 31) * TS;
  for (int k = 0; k < N; ++k) {
    double v = (L.l < 32 && L.l >= k) ? __builtin_fabs(tA[k * TS + L.l]) : -1.0;
    int idx = L.l;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double ov = shfl(v, L.l ^ off);
      const int oi = __builtin_amdgcn_ds_bpermute((L.l ^ off) << 2, idx);
      if (ov > v || (ov == v && oi < idx)) {
        v = ov;
        idx = oi;
      }
    }
    // (NaN entries can leave the lanes disagreeing: keep p a row k..31 anyway)
    const int p = __builtin_amdgcn_readfirstlane(min(max(idx, k), N - 1));
    if (p != k) {
      const double ak = col[k], ap = col[p];
      col[k] = ap;
      col[p] = ak;
    }
    wave_sync();
    const double d = tA[k * TS + k];
    const double rk = col[k] / d;
    col[k] = rk;
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (8 * q + 7 <= k) continue;  // wave-uniform: no row of this chunk is below k
      double f[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) f[t] = tA[k * TS + 8 * q + t];
      wave_sync();
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int i = 8 * q + t;
        if (i > k) col[i] = fma_(-f[t], rk, col[i]);
      }
      wave_sync();
    }
  }
  // back substitution: row k of X^T is final once the columns above k are done
  for (int k = N - 1; k > 0; --k) {
    const double xk = col[k];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (8 * q >= k) continue;  // wave-uniform: no row of this chunk is above k
      double f[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) f[t] = tA[k * TS + 8 * q + t];
      wave_sync();
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int i = 8 * q + t;
        if (i < k) col[i] = fma_(-f[t], xk, col[i]);
      }
      wave_sync();
    }
  }
}

// D = T (I - P)^-1 for P in the tile t1 (t2: scratch), by LU
__device__ __forceinline__ void right_solve_tile(double* t1, double* t2, const Mat& T, Mat& D,
                                                 const Lane& L) {
  Mat A;
  get(t1, L, A);
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) A.v[b][r] = delta(L, b, r) - A.v[b][r];
  wave_sync();
  put_t(t1, L, A);
  put_t(t2, L, T);
  wave_sync();
  gauss_jordan(t1, t2, L);
  wave_sync();
  get_t(t2, L, D);
  wave_sync();
}

// row of register r in lane group h for the register layout of RowT: the FP64
// instruction's interleaved rows h + 4 r, the FP32 instruction's blocked 4 h + r
template <typename RowT>
__device__ __forceinline__ int rrow(int h, int r) {
  return sizeof(RowT) == 8 ? h + 4 * r : 4 * h + r;
}

// phase matrices and thin-layer start of a scattering layer
// (src/phase_matrix.hpp:61-181, src/doubling.hpp:83-147), entries in the
// register layout of RowT: T0 (SFORM: S0 = I - T0 = tau0 G++, formed without
// cancellation) and R0 of rows 16 (b & 1) + rrow(h, r), columns 16 (b >> 1) + j
struct Start32 {
  Mat R, T;
  Vec y, sus, sds;
  double gam, gk, tau0;
  int nn;
};
template <typename RowT, bool SFORM>
__device__ __forceinline__ void start32(const FwdArgs& a, long task, int w, int l, long gw,
                                        const Lane& L, double tau, double omega_in, int mode,
                                        Start32& st) {
  const int Lay = a.L, j = L.j, h = L.h;
  const double mu_c[2] = {a.mu[j], a.mu[16 + j]};
  const double wt_c[2] = {a.wt[j], a.wt[16 + j]};
  auto row = [&](int b, int r) { return bi(b) + rrow<RowT>(h, r); };
  auto dlt = [&](int b, int r) { return (row(b, r) == bj(b) + j) ? 1.0 : 0.0; };
  const double omega = omega_in > 1.0 ? 1.0 : omega_in;
  const double con = 2.0 * omega * kPi;
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

  // phase matrices (src/phase_matrix.hpp:73-91): P(i, c) = sum_l c_l P_l(mu_i) P_l(+-mu_c)
  Mat Ppp, Ppm;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Ppp.v[b][r] = 0.0;
      Ppm.v[b][r] = 0.0;
    }
  double sign = 1.0;
  double sp[2] = {0.0, 0.0}, smv[2] = {0.0, 0.0};  // solar phase vectors, elements j, 16 + j
  for (int ll = 0; ll < Lm; ++ll) {
    const double chil = (mode == 3) ? (chi[ll] - f) * inv1mf : chi[ll];
    const double cl = (2 * ll + 1) * chil;
    const double* Pl = a.Pl + ll * N;
    const double plc[2] = {Pl[j], Pl[16 + j]};
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double pi = cl * Pl[16 * I + rrow<RowT>(h, r)];
#pragma unroll
        for (int J = 0; J < 2; ++J) {
          const double term = pi * plc[J];
          Ppp.v[I + 2 * J][r] += term;
          Ppm.v[I + 2 * J][r] += sign * term;
        }
      }
    if (a.has_solar) {
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        const double term = cl * plc[J] * a.Pl0[ll];
        sp[J] += term;
        smv[J] += sign * term;
      }
    }
    sign = -sign;
  }
  {
    // Hansen column normalisation (src/phase_matrix.hpp:93-110)
    double part[2] = {0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Ppp.v[b][r] = Ppp.v[b][r] / (2.0 * kPi);
        Ppm.v[b][r] = Ppm.v[b][r] / (2.0 * kPi);
        part[b >> 1] += (Ppp.v[b][r] + Ppm.v[b][r]) * a.wt[bi(b) + rrow<RowT>(h, r)];
      }
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      const double csum = w16::col_reduce<double>(part[J], L);
      if (csum > 0.0) {
        const double corr = 1.0 / (2.0 * kPi * csum);
#pragma unroll
        for (int I = 0; I < 2; ++I)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            Ppp.v[I + 2 * J][r] *= corr;
            Ppm.v[I + 2 * J][r] *= corr;
          }
      }
    }
  }
  double pplus[2] = {0.0, 0.0}, pminus[2] = {0.0, 0.0};
  if (a.has_solar) {  // src/phase_matrix.hpp:153-180
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      pplus[J] = sp[J] / (2.0 * kPi);
      pminus[J] = smv[J] / (2.0 * kPi);
    }
    const double s = row_sum<double>((pplus[0] + pminus[0]) * wt_c[0] +
                                     (pplus[1] + pminus[1]) * wt_c[1]);
    if (s > 0.0) {
      const double corr = 1.0 / (2.0 * kPi * s);
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        pplus[J] *= corr;
        pminus[J] *= corr;
      }
    }
  }

  // thin-layer initialisation (src/doubling.hpp:83-147)
  const int nn0 = (int)(log(tau) / log(2.0)) + ipow0(omega);
  const int nn = nn0 < 1 ? 1 : nn0;
  const double tau0 = tau * ldexp(1.0, -nn);
  Mat& Rm = st.R;
  Mat& Tm = st.T;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double mui = a.mu[bi(b) + rrow<RowT>(h, r)];
      const double wc = wt_c[b >> 1];
      const double d = dlt(b, r);
      const double gpp = (d + (-con) * (Ppp.v[b][r] * wc)) / mui;
      const double gpm = con * (Ppm.v[b][r] * wc) / mui;
      Tm.v[b][r] = SFORM ? tau0 * gpp : d - tau0 * gpp;
      Rm.v[b][r] = tau0 * gpm;
    }
  Vec& y = st.y;
  Vec& sus = st.sus;
  Vec& sds = st.sds;
  double& gam = st.gam;
  y = Vec{(1.0 - omega) * tau0 / mu_c[0], (1.0 - omega) * tau0 / mu_c[1]};
  sus = Vec{0.0, 0.0};
  sds = Vec{0.0, 0.0};
  gam = 0.0;
  if (a.has_solar) {
    const double Ftop = a.F0 * exp(-a.tau_cum[(long)w * (Lay + 1) + l] / a.mu0);
    const double b0 = omega * tau0 / mu_c[0] * Ftop, b1 = omega * tau0 / mu_c[1] * Ftop;
    sus = Vec{b0 * pminus[0], b1 * pminus[1]};
    sds = Vec{b0 * pplus[0], b1 * pplus[1]};
    gam = exp(-tau0 / a.mu0);
  }
  st.gk = 0.5 * tau0;
  st.tau0 = tau0;
  st.nn = nn;
}

// ---------------------------------------------------------------------------
// layer: phase matrices + doubling (src/phase_matrix.hpp:61-181,
// src/doubling.hpp:29-230); one wave per (wavenumber, layer).
// REDO: only the layers the mixed-precision kernel leaves to FP64 (the routed
// scattering layers, and the transparent and purely absorbing ones)
// ---------------------------------------------------------------------------
// 2 waves per SIMD (256 VGPRs, some spilled) measured 27 % faster than the
// spill-free single wave
template <bool REDO>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2, 8)))
layer32_kernel(FwdArgs a) {
  __shared__ double smem[WAVES * SLOT];
  const int wv = wave_index();
  const long task = (long)blockIdx.x * WAVES + wv;
  const int Lay = a.L;
  if (task >= (long)a.W * Lay) return;
  if constexpr (REDO) {
    if (!a.redo[task] && a.tau_s[task] > 0.0 && a.omega_s[task] > 0.0) return;
  }
  const Lane L;
  const int j = L.j, h = L.h;
  const int w = (int)(task / Lay), l = (int)(task % Lay);
  const long gw = (long)(a.w0 + w);
  double* sm = smem + wv * SLOT;
  double* tile = sm;
  double* tile2 = sm + TILE;
  double* v0 = sm + 2 * TILE;

  constexpr int REC = 2 * N * N + 2 * N;
  double* out = a.lay + task * REC;
  auto store = [&](double* m, const Mat& x) {  // column-major record
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) m[(bj(b) + j) * N + bi(b) + rw(h, r)] = x.v[b][r];
  };

  const double tau = a.tau_s[task];
  const double omega_in = a.omega_s[task];
  const int mode = a.mode[task];
  const double Btop = a.B ? a.B[gw * (Lay + 1) + lev(a, l)] : 0.0;
  const double Bbot = a.B ? a.B[gw * (Lay + 1) + lev(a, l + 1)] : 0.0;
  const double B_bar = (Bbot + Btop) / 2.0;
  const double B_d = (tau > 0.0) ? (Bbot - Btop) / tau : 0.0;
  const double mu_c[2] = {a.mu[j], a.mu[16 + j]};

  if (tau <= 0.0) {  // transparent layer (src/doubling.hpp:52-53)
    Mat Z, I1;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Z.v[b][r] = 0.0;
        I1.v[b][r] = delta(L, b, r);
      }
    store(out, Z);
    store(out + N * N, I1);
    if (h == 0)
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        out[2 * N * N + 16 * J + j] = 0.0;
        out[2 * N * N + N + 16 * J + j] = 0.0;
        if (a.basis) {
          a.basis[task * 2 * N + 16 * J + j] = 0.0;
          a.basis[task * 2 * N + N + 16 * J + j] = 0.0;
        }
      }
    if (j == 0 && h == 0) a.lflag[task] = 0;
    return;
  }
  if (omega_in <= 0.0) {  // pure absorption (src/doubling.hpp:56-77)
    double trans[2], omt[2], slope[2];
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      const double tex = -tau / mu_c[J];
      trans[J] = (tex > -200.0) ? exp(tex) : 0.0;
      omt[J] = 1.0 - trans[J];
      slope[J] = mu_c[J] * omt[J] - 0.5 * tau * (1.0 + trans[J]);
    }
    Mat Z, Tr;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Z.v[b][r] = 0.0;
        Tr.v[b][r] = delta(L, b, r) * trans[b >> 1];
      }
    store(out, Z);
    store(out + N * N, Tr);
    if (h == 0)
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        out[2 * N * N + 16 * J + j] = B_bar * omt[J] + B_d * slope[J];
        out[2 * N * N + N + 16 * J + j] = B_bar * omt[J] - B_d * slope[J];
        if (a.basis) {
          a.basis[task * 2 * N + 16 * J + j] = omt[J];
          a.basis[task * 2 * N + N + 16 * J + j] = slope[J];
        }
      }
    if (j == 0 && h == 0) a.lflag[task] = 0;
    return;
  }

  // ---- scattering layer ----
  Start32 st;
  start32<double, false>(a, task, w, l, gw, L, tau, omega_in, mode, st);
  const int nn = st.nn;
  Mat Rm = st.R, Tm = st.T;
  Vec y = st.y, z{0.0, 0.0}, sus = st.sus, sds = st.sds;
  double gam = st.gam, gk = st.gk;

  double* v1 = v0 + N;
  double* v2 = v1 + N;
  double* v3 = v2 + N;

  // doubling recursion (src/doubling.hpp:149-209)
  for (int k = 0; k < nn; ++k) {
    put(tile, L, Rm);
    put_vec(v0, L, sds);
    put_vec(v1, L, sus);
    wave_sync();
    Mat P;
    mul(tile, Rm, L, P);  // R^2
    Vec b12[2] = {Vec{0.0, 0.0}, Vec{0.0, 0.0}};
    if (a.has_solar) matvecs<2>(tile, v0, L, b12);  // R s_down_sol, R s_up_sol
    wave_sync();
    put(tile, L, P);
    wave_sync();
    Mat X;
    if (inverse_ns(P, tile, tile2, X, L)) {  // (I - R^2)^-1 ...
      right_solve_tile(tile, tile2, Tm, X, L);  // ... or TG by a solve
    } else {
      put(tile2, L, Tm);
      wave_sync();
      mul(tile2, X, L, X);  // TG = T (I - R^2)^-1
    }
    wave_sync();
    put(tile2, L, X);
    const Vec zpgy{z.lo + gk * y.lo, z.hi + gk * y.hi};
    put_vec(v0, L, zpgy);
    put_vec(v1, L, y);
    put_vec(v2, L, Vec{b12[0].lo + gam * sus.lo, b12[0].hi + gam * sus.hi});
    put_vec(v3, L, Vec{gam * b12[1].lo + sds.lo, gam * b12[1].hi + sds.hi});
    wave_sync();
    mul(tile2, Rm, L, X);  // TGR = TG R
    mul(tile2, Tm, L, P);  // T_new = TG T
    Vec s4[4];
    if (a.has_solar) {
      matvecs<4>(tile2, v0, L, s4);
    } else {
      Vec s2[2];
      matvecs<2>(tile2, v0, L, s2);
      s4[0] = s2[0];
      s4[1] = s2[1];
      s4[2] = Vec{0.0, 0.0};
      s4[3] = Vec{0.0, 0.0};
    }
    wave_sync();
    put(tile, L, X);
    wave_sync();
    mul_add(tile, Tm, Rm, L, Rm);  // R_new = R + TGR T
    Vec s2[2];
    matvecs<2>(tile, v0, L, s2);  // TGR (z + g y), TGR y
    wave_sync();
    Tm = P;
    const Vec znew{(s4[0].lo - s2[0].lo) + z.lo - gk * y.lo,
                   (s4[0].hi - s2[0].hi) + z.hi - gk * y.hi};
    const Vec ynew{s4[1].lo + s2[1].lo + y.lo, s4[1].hi + s2[1].hi + y.hi};
    z = znew;
    y = ynew;
    if (a.has_solar) {
      sus = Vec{s4[2].lo + sus.lo, s4[2].hi + sus.hi};
      sds = Vec{s4[3].lo + gam * sds.lo, s4[3].hi + gam * sds.hi};
      gam = gam * gam;
    }
    gk = 2.0 * gk;
  }

  // layer assembly (src/doubling.hpp:211-229), sources combined
  store(out, Rm);
  store(out + N * N, Tm);
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

// ---------------------------------------------------------------------------
// adding sweeps + interfaces (src/adding.hpp:69-258, src/solver.cpp:516-700),
// one wave per wavenumber; the two inverses of each general add and of each
// interface by inverse_ns, with a solve against the right-hand side where
// Newton-Schulz does not apply
// ---------------------------------------------------------------------------

// x = (I - P)^-1 b for P in the tile t1 (t2: scratch), b in the LDS slot vb,
// by LU: x^T (I - P)^T = b^T is X A = T with A^T = I - P and b in
// the first column of T^T
__device__ __forceinline__ Vec left_solve_tile(double* t1, double* t2, const double* vb,
                                               const Lane& L) {
  Mat A;
  get(t1, L, A);
  const double b = (L.l < 32) ? vb[L.l] : 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) A.v[q][r] = delta(L, q, r) - A.v[q][r];
  wave_sync();
  put(t1, L, A);
  if (L.l < 32) t2[L.l] = b;  // column 0 of T^T; the other columns are not read back
  wave_sync();
  gauss_jordan(t1, t2, L);
  wave_sync();
  const Vec x{t2[L.j], t2[16 + L.j]};
  wave_sync();
  return x;
}

// running composite: R_ab, R_ba, T_ab, T_ba; s_up, s_down
struct Comp {
  Mat Rab, Rba, Tab, Tba;
  Vec su, sd;
  int scat;
};

struct Ws {
  double *tile, *tile2, *v0, *v1, *v2, *v3;
};

// column-major 32 x 32 in global memory
__device__ __forceinline__ void gload(const double* g, const Lane& L, Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m.v[b][r] = g[(bj(b) + L.j) * N + bi(b) + rw(L.h, r)];
}
__device__ __forceinline__ void gstore(double* g, const Lane& L, const Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) g[(bj(b) + L.j) * N + bi(b) + rw(L.h, r)] = m.v[b][r];
}
// elements j, 16 + j of a vector in global memory (lanes of group 0)
__device__ __forceinline__ void store_vec(double* g, const Lane& L, const Vec& x) {
  if (L.h == 0) {
    g[L.j] = x.lo;
    g[16 + L.j] = x.hi;
  }
}

// the transpose of a column-major 32 x 32 in global memory
__device__ __forceinline__ void gload_t(const double* g, const Lane& L, Mat& m) {
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) m.v[b][r] = g[(bi(b) + rw(L.h, r)) * N + bj(b) + L.j];
}
__device__ __forceinline__ Vec gvec(const double* g, const Lane& L) {
  return Vec{g[L.j], g[16 + L.j]};
}
// diagonal M(c, c) for c = j, 16 + j
__device__ __forceinline__ Vec diag(const Mat& m, const Lane& L) {
  w16::Mat<double> b0, b3;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    b0.v[r] = m.v[0][r];
    b3.v[r] = m.v[3][r];
  }
  return Vec{w16::diag<double>(b0, L), w16::diag<double>(b3, L)};
}
// this lane's rows of a vector in an LDS slot: t[b][r] = v(16 (b & 1) + h + 4r)
__device__ __forceinline__ void rows_of(const double* v, const Lane& L, double (&t)[2][4]) {
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int r = 0; r < 4; ++r) t[I][r] = v[16 * I + rw(L.h, r)];
}
__device__ __forceinline__ double col_of(const Vec& v, int b) { return (b >> 1) ? v.hi : v.lo; }
__device__ __forceinline__ Vec vadd(const Vec& a, const Vec& b) {
  return Vec{a.lo + b.lo, a.hi + b.hi};
}
__device__ __forceinline__ Vec vmul(const Vec& a, const Vec& b) {
  return Vec{a.lo * b.lo, a.hi * b.hi};
}
__device__ __forceinline__ Vec matvec(const double* X, const double* v, const Lane& L) {
  Vec s[1];
  matvecs<1>(X, v, L, s);
  return s[0];
}

// D = T (I - P)^-1, P in registers and in tile s.tile, T in registers
__device__ __forceinline__ void right_inverse(const Mat& P, const Mat& T, Mat& D, const Ws& s,
                                              const Lane& L) {
  if (inverse_ns(P, s.tile, s.tile2, D, L)) {
    right_solve_tile(s.tile, s.tile2, T, D, L);
  } else {
    put(s.tile2, L, T);
    wave_sync();
    mul(s.tile2, D, L, D);
  }
  wave_sync();
}

// (I - P)^-1 b, P in registers and in tile s.tile, b in the LDS slot vb
__device__ __forceinline__ Vec left_inverse(const Mat& P, const double* vb, const Ws& s,
                                            const Lane& L) {
  Mat X;
  Vec x;
  if (inverse_ns(P, s.tile, s.tile2, X, L)) {
    x = left_solve_tile(s.tile, s.tile2, vb, L);
  } else {
    put(s.tile2, L, X);
    wave_sync();
    x = matvec(s.tile2, vb, L);
  }
  wave_sync();
  return x;
}

// cotangents of I_up(l) and I_dn(l) (elements j, 16 + j) from the output
// cotangents, as the N = 16 reverse takes them (w16::ibar16)
__device__ __forceinline__ void ibar32(const FwdArgs& a, long gw, int l, const Lane& L, Vec& up,
                                       Vec& dn) {
  w16::ibar16(a, gw, l, a.mu[L.j], a.wt[L.j], up.lo, dn.lo);
  w16::ibar16(a, gw, l, a.mu[16 + L.j], a.wt[16 + L.j], up.hi, dn.hi);
}

// R or T of a layer record, or (ADJ) its transpose
template <bool ADJ>
__device__ __forceinline__ void rec_mat(const double* m, const Lane& L, Mat& M) {
  if constexpr (ADJ) gload_t(m, L, M);
  else gload(m, L, M);
}

// rbase step: new = add(top = stored layer, bot = C)  (src/adding.hpp:246-258)
// (ADJ: the layer of the transposed system -- transposed R and T, the
// cotangent sources tsu, tsd)
template <bool ADJ>
__device__ __forceinline__ void add32_on_top(Comp& C, const double* rec, int lscat, const Vec& tsu,
                                             const Vec& tsd, const Lane& L, const Ws& s) {
  const double* Rt = rec;
  const double* Tt = rec + N * N;
  Mat A, B, X, D1;
  if (!lscat) {
    // NonscatTop (src/adding.hpp:117-174): t = diag(top.T_ab)
    const Vec t{Tt[L.j * N + L.j], Tt[(16 + L.j) * N + 16 + L.j]};
    put_vec(s.v0, L, t);
    put_vec(s.v1, L, tsd);
    put(s.tile, L, C.Rab);
    put(s.tile2, L, C.Tab);
    wave_sync();
    const Vec rbc = matvec(s.tile, s.v1, L);   // bot.R_ab s1_down
    const Vec tbc = matvec(s.tile2, s.v1, L);  // bot.T_ab s1_down
    double tv[2][4];
    rows_of(s.v0, L, tv);
    wave_sync();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double ti = tv[b & 1][r], tc = col_of(t, b);
        C.Rab.v[b][r] = ti * C.Rab.v[b][r] * tc;
        C.Tab.v[b][r] = C.Tab.v[b][r] * tc;
        C.Tba.v[b][r] = ti * C.Tba.v[b][r];
      }
    const Vec nsu = vadd(tsu, vmul(t, vadd(C.su, rbc)));
    const Vec nsd = vadd(C.sd, tbc);
    C.su = nsu;
    C.sd = nsd;
    return;  // scat stays bot's
  }
  if (!C.scat) {
    // NonscatBot (src/adding.hpp:181-239): t = diag(bot.T_ab)
    const Vec t = diag(C.Tab, L);
    put_vec(s.v0, L, t);
    put_vec(s.v1, L, C.su);
    rec_mat<ADJ>(Tt, L, B);
    rec_mat<ADJ>(Rt, L, A);
    put(s.tile, L, B);
    put(s.tile2, L, A);
    wave_sync();
    const Vec tba = matvec(s.tile, s.v1, L);   // top.T_ba s2_up
    const Vec rba = matvec(s.tile2, s.v1, L);  // top.R_ba s2_up
    double tv[2][4];
    rows_of(s.v0, L, tv);
    wave_sync();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double ti = tv[b & 1][r], tc = col_of(t, b);
        C.Rab.v[b][r] = A.v[b][r];
        C.Rba.v[b][r] = ti * A.v[b][r] * tc;
        C.Tab.v[b][r] = ti * B.v[b][r];
        C.Tba.v[b][r] = B.v[b][r] * tc;
      }
    const Vec nsu = vadd(tsu, tba);
    const Vec nsd = vadd(C.sd, vmul(t, vadd(tsd, rba)));
    C.su = nsu;
    C.sd = nsd;
    C.scat = 1;
    return;
  }
  // General (src/adding.hpp:69-110)
  rec_mat<ADJ>(Rt, L, A);
  put(s.tile, L, C.Rab);
  put(s.tile2, L, A);
  put_vec(s.v0, L, tsd);
  put_vec(s.v1, L, C.su);
  wave_sync();
  mul(s.tile, A, L, X);                                  // P1 = bot.R_ab top.R_ba
  const Vec rhs_up = vadd(C.su, matvec(s.tile, s.v0, L));  // s2_up + R_bc s1_down
  const Vec rhs_dn = vadd(tsd, matvec(s.tile2, s.v1, L));  // s1_down + R_ba s2_up
  wave_sync();
  put(s.tile, L, X);
  wave_sync();
  rec_mat<ADJ>(Tt, L, B);
  right_inverse(X, B, D1, s, L);                         // D1 = top.T_ba A1^-1
  rec_mat<ADJ>(Rt, L, A);
  put(s.tile2, L, A);
  wave_sync();
  mul(s.tile2, C.Rab, L, X);                             // P2 = top.R_ba bot.R_ab
  wave_sync();
  put(s.tile, L, X);
  wave_sync();
  right_inverse(X, C.Tab, B, s, L);                      // D2 = bot.T_ab A2^-1
  put_vec(s.v2, L, rhs_up);
  put_vec(s.v3, L, rhs_dn);
  put(s.tile, L, B);   // D2
  put(s.tile2, L, D1);  // D1
  wave_sync();
  const Vec nsd = vadd(C.sd, matvec(s.tile, s.v3, L));
  const Vec nsu = vadd(tsu, matvec(s.tile2, s.v2, L));
  rec_mat<ADJ>(Rt, L, A);
  mul(s.tile, A, L, X);                                  // D2 top.R_ba
  rec_mat<ADJ>(Tt, L, A);
  mul(s.tile, A, L, C.Tab);                              // T_ab = D2 top.T_ab
  mul(s.tile2, C.Rab, L, B);                             // D1 bot.R_ab
  mul(s.tile2, C.Tba, L, D1);                            // new T_ba = D1 bot.T_ba
  wave_sync();
  put(s.tile, L, X);
  put(s.tile2, L, B);
  wave_sync();
  mul_add(s.tile, C.Tba, C.Rba, L, C.Rba);               // R_ba = bot.R_ba + D2 R_ba T_ba
  rec_mat<ADJ>(Rt, L, B);
  mul_add(s.tile2, A, B, L, C.Rab);                      // R_ab = top.R_ab + D1 R_ab T_ab
  wave_sync();
  C.Tba = D1;
  C.su = nsu;
  C.sd = nsd;
  C.scat = 1;
}

// rtop step: new = add(top = C, bot = stored layer)
// (ADJ: as add32_on_top)
template <bool ADJ>
__device__ __forceinline__ void add32_below(Comp& C, const double* rec, int lscat, const Vec& bsu,
                                            const Vec& bsd, const Lane& L, const Ws& s) {
  const double* Rb = rec;
  const double* Tb = rec + N * N;
  Mat A, B, X, D1;
  if (!C.scat) {
    // NonscatTop with top = C: t = diag(C.T_ab)
    const Vec t = diag(C.Tab, L);
    put_vec(s.v0, L, t);
    put_vec(s.v1, L, C.sd);
    rec_mat<ADJ>(Rb, L, A);
    rec_mat<ADJ>(Tb, L, B);
    put(s.tile, L, A);
    put(s.tile2, L, B);
    wave_sync();
    const Vec rbc = matvec(s.tile, s.v1, L);   // bot.R_ab s1_down
    const Vec tbc = matvec(s.tile2, s.v1, L);  // bot.T_ab s1_down
    double tv[2][4];
    rows_of(s.v0, L, tv);
    wave_sync();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double ti = tv[b & 1][r], tc = col_of(t, b);
        C.Rab.v[b][r] = ti * A.v[b][r] * tc;
        C.Rba.v[b][r] = A.v[b][r];
        C.Tab.v[b][r] = B.v[b][r] * tc;
        C.Tba.v[b][r] = ti * B.v[b][r];
      }
    const Vec nsu = vadd(C.su, vmul(t, vadd(bsu, rbc)));
    const Vec nsd = vadd(bsd, tbc);
    C.su = nsu;
    C.sd = nsd;
    C.scat = lscat;
    return;
  }
  if (!lscat) {
    // NonscatBot with bot = layer: t = diag(layer T)
    const Vec t{Tb[L.j * N + L.j], Tb[(16 + L.j) * N + 16 + L.j]};
    put_vec(s.v0, L, t);
    put_vec(s.v1, L, bsu);
    put(s.tile, L, C.Tba);
    put(s.tile2, L, C.Rba);
    wave_sync();
    const Vec tba = matvec(s.tile, s.v1, L);   // top.T_ba s2_up
    const Vec rba = matvec(s.tile2, s.v1, L);  // top.R_ba s2_up
    double tv[2][4];
    rows_of(s.v0, L, tv);
    wave_sync();
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double ti = tv[b & 1][r], tc = col_of(t, b);
        C.Rba.v[b][r] = ti * C.Rba.v[b][r] * tc;
        C.Tab.v[b][r] = ti * C.Tab.v[b][r];
        C.Tba.v[b][r] = C.Tba.v[b][r] * tc;
      }
    const Vec nsu = vadd(C.su, tba);
    const Vec nsd = vadd(bsd, vmul(t, vadd(C.sd, rba)));
    C.su = nsu;
    C.sd = nsd;
    return;
  }
  // General with top = C, bot = layer
  rec_mat<ADJ>(Rb, L, A);
  put(s.tile, L, A);
  put(s.tile2, L, C.Rba);
  put_vec(s.v0, L, C.sd);
  put_vec(s.v1, L, bsu);
  wave_sync();
  mul(s.tile, C.Rba, L, X);                               // P1 = bot.R_ab top.R_ba
  const Vec rhs_up = vadd(bsu, matvec(s.tile, s.v0, L));   // s2_up + R_bc s1_down
  const Vec rhs_dn = vadd(C.sd, matvec(s.tile2, s.v1, L));  // s1_down + R_ba s2_up
  wave_sync();
  put(s.tile, L, X);
  wave_sync();
  right_inverse(X, C.Tba, D1, s, L);                      // D1 = top.T_ba A1^-1
  put(s.tile2, L, C.Rba);
  rec_mat<ADJ>(Rb, L, A);
  wave_sync();
  mul(s.tile2, A, L, X);                                  // P2 = top.R_ba bot.R_ab
  wave_sync();
  put(s.tile, L, X);
  wave_sync();
  rec_mat<ADJ>(Tb, L, B);
  right_inverse(X, B, A, s, L);                           // D2 = bot.T_ab A2^-1
  put_vec(s.v2, L, rhs_up);
  put_vec(s.v3, L, rhs_dn);
  put(s.tile, L, D1);
  put(s.tile2, L, A);  // D2
  wave_sync();
  const Vec nsu = vadd(C.su, matvec(s.tile, s.v2, L));
  const Vec nsd = vadd(bsd, matvec(s.tile2, s.v3, L));
  rec_mat<ADJ>(Rb, L, B);
  mul(s.tile, B, L, X);                                   // D1 bot.R_ab
  rec_mat<ADJ>(Tb, L, B);
  mul(s.tile, B, L, D1);                                  // new T_ba = D1 bot.T_ba
  mul(s.tile2, C.Rba, L, A);                              // D2 top.R_ba
  wave_sync();
  put(s.tile, L, X);
  wave_sync();
  mul_add(s.tile, C.Tab, C.Rab, L, C.Rab);                // R_ab = top.R_ab + D1 R_ab T_ab
  mul(s.tile2, C.Tab, L, X);                              // new T_ab = D2 top.T_ab
  wave_sync();
  put(s.tile2, L, A);
  wave_sync();
  rec_mat<ADJ>(Rb, L, A);
  mul_add(s.tile2, B, A, L, C.Rba);                       // R_ba = bot.R_ba + D2 R_ba T_ba
  wave_sync();
  C.Tab = X;
  C.Tba = D1;
  C.su = nsu;
  C.sd = nsd;
  C.scat = 1;
}

// The reference's form, for the columns the prep kernel flags (consecutive
// scattering layers with unstable doubling starts; the N = 16 adding's
// composite kernel gives the reason); adding32_kernel below does the others.
// 2 waves per SIMD measured 16 % faster than one wave with all 512 registers.
// ADJ: the transposed system of the reverse pass for these columns (as
// adding16_composite_kernel<true>)
template <bool ADJ>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2, 8)))
adding32_composite_kernel(FwdArgs a) {
  __shared__ double smem[WAVES * SLOT];
  const int wv = wave_index();
  int w = blockIdx.x * WAVES + wv;
  if (!composite_column(a, w)) return;
  const Lane L;
  const int j = L.j;
  const long gw = (long)(a.w0 + w);
  const int Lay = a.L;
  double* sm = smem + wv * SLOT;
  const Ws s{sm, sm + TILE, sm + 2 * TILE, sm + 2 * TILE + N, sm + 2 * TILE + 2 * N,
             sm + 2 * TILE + 3 * N};
  constexpr int REC = 2 * N * N + 2 * N;
  constexpr int RBR = 2 * N * N + N;
  const double* lay = a.lay + (long)w * Lay * REC;
  const int* lflag = a.lflag + (long)w * Lay;
  // the column's region of the adding store is L (2N^2 + 4N) doubles, as the
  // base sweep's (RBR32 per layer): this kernel runs beside adding32_kernel, so
  // its columns must not reach into that kernel's columns
  double* rbw = a.rb + (long)w * Lay * (2 * N * N + 4 * N);
  const double mu_c[2] = {a.mu[j], a.mu[16 + j]};
  const double wt_c[2] = {a.wt[j], a.wt[16 + j]};
  const double* Bw = a.B ? a.B + gw * (Lay + 1) : nullptr;
  auto Blev = [&](int l) -> double { return Bw ? Bw[lev(a, l)] : 0.0; };
  // the sources of layer k, or (ADJ) the cotangents of I_dn(k) and I_up(k+1)
  auto sources = [&](int k, Vec& su, Vec& sd) {
    if constexpr (ADJ) {
      Vec u, d;
      ibar32(a, gw, k, L, u, d);
      su = d;
      ibar32(a, gw, k + 1, L, u, d);
      sd = u;
    } else {
      const double* rec = lay + (long)k * REC;
      su = gvec(rec + 2 * N * N, L);
      sd = gvec(rec + 2 * N * N + N, L);
    }
  };

  // boundary conditions (src/solver.cpp:385-419, 516-551, 585-603)
  const double Bs = a.surf_w ? a.surf_w[gw] : a.surf_em;
  const double Btop = a.top_w ? a.top_w[gw] : a.top_em;
  const double A = a.albedo;
  const int has_surface = !a.diffusion && (A > 0.0 || Bs > 0.0);
  const int ltot = Lay + (has_surface ? 1 : 0);
  const double tau_tot = a.tau_cum[(long)w * (Lay + 1) + Lay];
  Vec Itd{Btop, Btop};
  Vec Ibu{0.0, 0.0};
  if constexpr (ADJ) {
    Vec u, d;
    ibar32(a, gw, 0, L, u, d);
    Itd = u;
    ibar32(a, gw, Lay, L, u, d);
    if (!has_surface) Ibu = d;
  } else if (a.diffusion) {
    const double Bb = Blev(Lay);
    const double dt = a.tau_s[(long)w * Lay + Lay - 1];
    const double dBdt = (dt > 0.0) ? (Bb - Blev(Lay - 1)) / dt : 0.0;
    Ibu = Vec{Bb + mu_c[0] * dBdt, Bb + mu_c[1] * dBdt};
  } else if (!has_surface) {
    Ibu = Vec{Bs, Bs};
  }

  Comp C;
  // ---- bottom-up sweep (src/solver.cpp:560-572) ----
  if (has_surface) {
    const double rr[2] = {2.0 * A * mu_c[0] * wt_c[0] * a.xfac, 2.0 * A * mu_c[1] * wt_c[1] * a.xfac};
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = bi(b) + rw(L.h, r), c = bj(b) + j;
        if constexpr (ADJ)  // R_s^T(i, c) = R_s(c, i)
          C.Rab.v[b][r] = (c < a.n_real) ? 2.0 * A * a.mu[i] * a.wt[i] * a.xfac : 0.0;
        else
          C.Rab.v[b][r] = (i < a.n_real) ? rr[b >> 1] : 0.0;
        C.Rba.v[b][r] = C.Rab.v[b][r];
        C.Tab.v[b][r] = 0.0;
        C.Tba.v[b][r] = 0.0;
      }
    if constexpr (ADJ) {
      Vec u, d;
      ibar32(a, gw, Lay, L, u, d);
      C.su = d;
    } else {
      double ssol = 0.0;
      if (a.has_solar && A > 0.0) ssol = (A / kPi) * a.F0 * a.mu0 * exp(-tau_tot / a.mu0);
      const double su = (1.0 - A) * Bs + ssol;
      C.su = Vec{su, su};
    }
    C.sd = Vec{0.0, 0.0};
    C.scat = A > 0.0;
  } else {
    const double* rec = lay + (long)(Lay - 1) * REC;
    rec_mat<ADJ>(rec, L, C.Rab);
    rec_mat<ADJ>(rec + N * N, L, C.Tab);
    C.Rba = C.Rab;
    C.Tba = C.Tab;
    sources(Lay - 1, C.su, C.sd);
    C.scat = lflag[Lay - 1];
  }
  for (int n = 1;; ++n) {  // C == rbase[n]
    if (n < ltot) {        // keep base composites n = 1..ltot-1 for the interfaces
      double* o = rbw + (long)(n - 1) * RBR;
      gstore(o, L, C.Rab);
      gstore(o + N * N, L, C.Tba);
      if (L.h == 0) {
        o[2 * N * N + j] = C.su.lo;
        o[2 * N * N + 16 + j] = C.su.hi;
      }
    }
    if (n == ltot) break;
    const int k = ltot - 1 - n;
    Vec su, sd;
    sources(k, su, sd);
    add32_on_top<ADJ>(C, lay + (long)k * REC, lflag[k], su, sd, L, s);
  }

  // interface intensities (the reverse pass's) or (ADJ) adjoint vectors
  double* vec = ADJ ? a.Av : a.Iv;
  auto emit = [&](int l, const Vec& Iup, const Vec& Idn) {
    if (vec && L.h == 0) {
      double* o = vec + ((long)w * (Lay + 1) + l) * 2 * N;
      store_vec(o, L, Iup);
      store_vec(o + N, L, Idn);
    }
    if constexpr (ADJ) return;
    const double c0 = 2.0 * kPi * wt_c[0] * mu_c[0], c1 = 2.0 * kPi * wt_c[1] * mu_c[1];
    const double fu = row_sum<double>(c0 * Iup.lo + c1 * Iup.hi);
    const double fdn = row_sum<double>(c0 * Idn.lo + c1 * Idn.hi);
    double J = row_sum<double>(0.5 * wt_c[0] * (Iup.lo + Idn.lo) +
                               0.5 * wt_c[1] * (Iup.hi + Idn.hi));
    if (j == 0 && L.h == 0) {
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
  };

  // ---- TOA (src/solver.cpp:628-642) ----
  {
    put(s.tile, L, C.Rab);
    put(s.tile2, L, C.Tba);
    put_vec(s.v0, L, Itd);
    put_vec(s.v1, L, Ibu);
    wave_sync();
    const Vec rI = matvec(s.tile, s.v0, L);
    const Vec tI = matvec(s.tile2, s.v1, L);
    wave_sync();
    emit(0, vadd(rI, vadd(tI, C.su)), Itd);
  }

  // ---- top-down sweep with interfaces (src/solver.cpp:574-583, 644-678) ----
  rec_mat<ADJ>(lay, L, C.Rab);
  rec_mat<ADJ>(lay + N * N, L, C.Tab);
  C.Rba = C.Rab;
  C.Tba = C.Tab;
  sources(0, C.su, C.sd);
  C.scat = lflag[0];
  for (int l = 1; l <= Lay; ++l) {  // C == rtop[l]
    const int n_base = ltot - l;
    Vec Iup, Idn;
    if (n_base > 0) {
      const double* o = rbw + (long)(n_base - 1) * RBR;
      const Vec sbu = gvec(o + 2 * N * N, L);
      Mat Bt, P;
      gload(o + N * N, L, Bt);
      put(s.tile, L, C.Tab);
      put(s.tile2, L, Bt);
      put_vec(s.v0, L, Itd);
      put_vec(s.v1, L, Ibu);
      wave_sync();
      const Vec tti = matvec(s.tile, s.v0, L);   // top.T_ab I_top_down
      const Vec tbi = matvec(s.tile2, s.v1, L);  // base.T_ba I_bot_up
      wave_sync();
      // M_up = I - base.R_ab top.R_ba; terms 2 and 3 (src/solver.cpp:24-46)
      gload(o, L, Bt);  // base.R_ab
      put(s.tile, L, Bt);
      put_vec(s.v0, L, tti);
      put_vec(s.v1, L, C.sd);
      wave_sync();
      mul(s.tile, C.Rba, L, P);
      Vec t23[2];
      matvecs<2>(s.tile, s.v0, L, t23);
      wave_sync();
      put(s.tile, L, P);
      put_vec(s.v2, L, vadd(vadd(vadd(tbi, t23[0]), t23[1]), sbu));
      wave_sync();
      Iup = left_inverse(P, s.v2, s, L);
      // M_dn = I - top.R_ba base.R_ab (src/solver.cpp:51-73)
      put(s.tile, L, C.Rba);
      put_vec(s.v0, L, tbi);
      put_vec(s.v1, L, sbu);
      wave_sync();
      mul(s.tile, Bt, L, P);
      Vec u23[2];
      matvecs<2>(s.tile, s.v0, L, u23);
      wave_sync();
      put(s.tile, L, P);
      put_vec(s.v2, L, vadd(vadd(vadd(tti, u23[0]), u23[1]), C.sd));
      wave_sync();
      Idn = left_inverse(P, s.v2, s, L);
    } else {
      put(s.tile, L, C.Tab);
      put(s.tile2, L, C.Rba);
      put_vec(s.v0, L, Itd);
      put_vec(s.v1, L, Ibu);
      wave_sync();
      const Vec tti = matvec(s.tile, s.v0, L);
      const Vec rbi = matvec(s.tile2, s.v1, L);
      wave_sync();
      Idn = vadd(tti, vadd(rbi, C.sd));
      Iup = Ibu;
    }
    emit(l, Iup, Idn);
    if (l < Lay) {
      Vec su, sd;
      sources(l, su, sd);
      add32_below<ADJ>(C, lay + (long)l * REC, lflag[l], su, sd, L, s);
    }
  }
}


// ---------------------------------------------------------------------------
// adding by the base sweep: the same system as the N = 16 adding16_kernel
// solves (see adrt_fwd16.hpp) -- one sweep of the base composites Rb(l), sb(l)
// from the bottom, one inverse and four products per scattering layer, then a
// downward sweep of matrix-vector products for the interface intensities
// ---------------------------------------------------------------------------
// per layer: G_l (column-major; the diagonal of T_l for a non-scattering
// layer), Rb(l+1) (column-major), d_l, sb(l+1), whether Rb(l+1) is non-zero
constexpr int RB_G = 0, RB_R = N * N, RB_D = 2 * N * N, RB_S = RB_D + N, RB_F = RB_S + N;
constexpr int RBR32 = 2 * N * N + 4 * N;
static_assert(RB_F < RBR32, "adding record fields past the record");

struct Base {
  Mat R;
  Vec s;
  int scat;
};

// G = (I - P)^-1 T for P in the tile t1 (t2: scratch), by LU: G^T (I - P)^T =
// T^T is X A = T' with A^T = I - P and T'^T = T, both column-major as they are
__device__ __forceinline__ void left_solve_mat_tile(double* t1, double* t2, const Mat& T, Mat& G,
                                                    const Lane& L) {
  Mat A;
  get(t1, L, A);
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) A.v[b][r] = delta(L, b, r) - A.v[b][r];
  wave_sync();
  put(t1, L, A);
  put(t2, L, T);
  wave_sync();
  gauss_jordan(t1, t2, L);
  wave_sync();
  get(t2, L, G);
  wave_sync();
}

// layer record rec on top of the base B: stores at o what the downward sweep
// needs at this layer, then replaces B by the base above the layer.  ADJ: the
// transposed system (transposed R and T; sources su, sd the cotangents of
// I_dn(l) and I_up(l+1))
template <bool ADJ>
__device__ __forceinline__ void base32_add(Base& B, const double* rec, int lscat, double* o,
                                           const Vec& su, const Vec& sd, const Lane& L,
                                           const Ws& s) {
  if (B.scat) gstore(o + RB_R, L, B.R);
  store_vec(o + RB_S, L, B.s);
  if (L.l == 0) o[RB_F] = B.scat ? 1.0 : 0.0;
  if (!lscat) {
    // T = diag(t), R = 0: G = T, d = s_dn, Rb = T Rb T, sb = s_up + T (Rb s_dn + sb)
    const double* Tt = rec + N * N;
    const Vec t{Tt[L.j * N + L.j], Tt[(16 + L.j) * N + 16 + L.j]};
    Vec rsd{0.0, 0.0};
    if (B.scat) {
      put_vec(s.v0, L, t);
      put_vec(s.v1, L, sd);
      put(s.tile, L, B.R);
      wave_sync();
      rsd = matvec(s.tile, s.v1, L);
      double tv[2][4];
      rows_of(s.v0, L, tv);
      wave_sync();
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) B.R.v[b][r] = tv[b & 1][r] * B.R.v[b][r] * col_of(t, b);
    }
    store_vec(o + RB_G, L, t);
    store_vec(o + RB_D, L, sd);
    B.s = vadd(su, vmul(t, vadd(rsd, B.s)));
    return;
  }
  if (!B.scat) {
    // nothing below reflects: G = T, d = R sb + s_dn, Rb = R, sb = s_up + T sb
    Mat T;
    rec_mat<ADJ>(rec, L, B.R);
    rec_mat<ADJ>(rec + N * N, L, T);
    put(s.tile, L, B.R);
    put(s.tile2, L, T);
    put_vec(s.v0, L, B.s);
    wave_sync();
    const Vec d = vadd(matvec(s.tile, s.v0, L), sd);
    const Vec ts = matvec(s.tile2, s.v0, L);
    wave_sync();
    gstore(o + RB_G, L, T);
    store_vec(o + RB_D, L, d);
    B.s = vadd(su, ts);
    B.scat = 1;
    return;
  }
  Mat A, P, G;
  rec_mat<ADJ>(rec, L, A);  // R_l
  put(s.tile, L, A);
  put_vec(s.v0, L, B.s);
  wave_sync();
  mul(s.tile, B.R, L, P);                           // P = R_l Rb(l+1)
  const Vec v = vadd(matvec(s.tile, s.v0, L), sd);  // R_l sb(l+1) + s_dn
  wave_sync();
  put(s.tile, L, P);
  put_vec(s.v1, L, v);
  wave_sync();
  Vec d;
  if (inverse_ns(P, s.tile, s.tile2, G, L)) {
    // I - P close to singular or the iteration slow: d and G by LU solves
    d = left_solve_tile(s.tile, s.tile2, s.v1, L);
    put(s.tile, L, P);
    rec_mat<ADJ>(rec + N * N, L, A);
    wave_sync();
    left_solve_mat_tile(s.tile, s.tile2, A, G, L);
  } else {
    put(s.tile2, L, G);
    rec_mat<ADJ>(rec + N * N, L, A);
    wave_sync();
    mul(s.tile2, A, L, G);          // G = X T_l
    d = matvec(s.tile2, s.v1, L);   // d = X (R_l sb + s_dn)
  }
  wave_sync();
  gstore(o + RB_G, L, G);
  store_vec(o + RB_D, L, d);
  put(s.tile, L, B.R);
  put_vec(s.v2, L, d);
  wave_sync();
  mul(s.tile, G, L, P);                             // Rb(l+1) G
  const Vec rd = matvec(s.tile, s.v2, L);           // Rb(l+1) d
  wave_sync();
  rec_mat<ADJ>(rec + N * N, L, A);
  put(s.tile, L, A);
  put_vec(s.v3, L, vadd(rd, B.s));
  wave_sync();
  rec_mat<ADJ>(rec, L, A);
  mul_add(s.tile, P, A, L, B.R);                    // Rb(l) = R_l + T_l Rb(l+1) G
  const Vec ts = matvec(s.tile, s.v3, L);
  wave_sync();
  B.s = vadd(su, ts);
}

// ADJ: the transposed system of the reverse pass (as adding16_kernel<true>);
// the columns the prep kernel flags take adding32_composite_kernel<ADJ>
template <bool ADJ>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(2, 8)))
adding32_kernel(FwdArgs a) {
  __shared__ double smem[WAVES * SLOT];
  const int wv = wave_index();
  const int w = blockIdx.x * WAVES + wv;
  if (w >= a.W || (a.ucol && a.ucol[w])) return;
  const Lane L;
  const int j = L.j;
  const long gw = (long)(a.w0 + w);
  const int Lay = a.L;
  double* sm = smem + wv * SLOT;
  const Ws s{sm, sm + TILE, sm + 2 * TILE, sm + 2 * TILE + N, sm + 2 * TILE + 2 * N,
             sm + 2 * TILE + 3 * N};
  constexpr int REC = 2 * N * N + 2 * N;
  const double* lay = a.lay + (long)w * Lay * REC;
  const int* lflag = a.lflag + (long)w * Lay;
  double* rbw = a.rb + (long)w * Lay * RBR32;
  const double mu_c[2] = {a.mu[j], a.mu[16 + j]};
  const double wt_c[2] = {a.wt[j], a.wt[16 + j]};
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
  Vec Itd{Btop, Btop};
  Vec Ibu{0.0, 0.0};
  if constexpr (ADJ) {
    // transposed system: the top boundary carries the cotangent of I_up(0),
    // a fixed I_up(L) the cotangent of I_dn(L)
    Vec u, d;
    ibar32(a, gw, 0, L, u, d);
    Itd = u;
    ibar32(a, gw, Lay, L, u, d);
    if (!has_surface) Ibu = d;
  } else if (a.diffusion) {
    const double Bb = Blev(Lay);
    const double dt = a.tau_s[(long)w * Lay + Lay - 1];
    const double dBdt = (dt > 0.0) ? (Bb - Blev(Lay - 1)) / dt : 0.0;
    Ibu = Vec{Bb + mu_c[0] * dBdt, Bb + mu_c[1] * dBdt};
  } else if (!has_surface) {
    Ibu = Vec{Bs, Bs};
  }

  // ---- the base below the last layer, then the base sweep ----
  Base B;
  if (has_surface) {
    const double rr[2] = {2.0 * A * mu_c[0] * wt_c[0] * a.xfac, 2.0 * A * mu_c[1] * wt_c[1] * a.xfac};
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = bi(b) + rw(L.h, r), c = bj(b) + j;
        if constexpr (ADJ)  // R_s^T(i, c) = R_s(c, i)
          B.R.v[b][r] = (c < a.n_real) ? 2.0 * A * a.mu[i] * a.wt[i] * a.xfac : 0.0;
        else
          B.R.v[b][r] = (i < a.n_real) ? rr[b >> 1] : 0.0;
      }
    if constexpr (ADJ) {
      Vec u, d;
      ibar32(a, gw, Lay, L, u, d);
      B.s = d;
    } else {
      double ssol = 0.0;
      if (a.has_solar && A > 0.0) ssol = (A / kPi) * a.F0 * a.mu0 * exp(-tau_tot / a.mu0);
      const double su = (1.0 - A) * Bs + ssol;
      B.s = Vec{su, su};
    }
    B.scat = A > 0.0;
  } else {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) B.R.v[b][r] = 0.0;
    B.s = Ibu;
    B.scat = 0;
  }
  for (int l = Lay - 1; l >= 0; --l) {
    const double* rec = lay + (long)l * REC;
    Vec su, sd;
    if constexpr (ADJ) {
      Vec u, d;
      ibar32(a, gw, l, L, u, d);
      su = d;
      ibar32(a, gw, l + 1, L, u, d);
      sd = u;
    } else {
      su = gvec(rec + 2 * N * N, L);
      sd = gvec(rec + 2 * N * N + N, L);
    }
    base32_add<ADJ>(B, rec, lflag[l], rbw + (long)l * RBR32, su, sd, L, s);
  }

  auto emit = [&](int l, const Vec& Iup, const Vec& Idn) {
    if (vec && L.h == 0) {
      store_vec(vec + (long)l * 2 * N, L, Iup);
      store_vec(vec + (long)l * 2 * N + N, L, Idn);
    }
    if constexpr (ADJ) return;
    const double c0 = 2.0 * kPi * wt_c[0] * mu_c[0], c1 = 2.0 * kPi * wt_c[1] * mu_c[1];
    const double fu = row_sum<double>(c0 * Iup.lo + c1 * Iup.hi);
    const double fdn = row_sum<double>(c0 * Idn.lo + c1 * Idn.hi);
    double J = row_sum<double>(0.5 * wt_c[0] * (Iup.lo + Idn.lo) +
                               0.5 * wt_c[1] * (Iup.hi + Idn.hi));
    if (j == 0 && L.h == 0) {
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
  };

  // ---- downward sweep (the interface intensities of src/solver.cpp:628-678) ----
  Vec Idn = Itd, Iup = B.s;
  if (B.scat) {
    put(s.tile, L, B.R);
    put_vec(s.v0, L, Itd);
    wave_sync();
    Iup = vadd(Iup, matvec(s.tile, s.v0, L));
    wave_sync();
  }
  emit(0, Iup, Idn);
  for (int l = 0; l < Lay; ++l) {
    const double* o = rbw + (long)l * RBR32;
    const Vec d = gvec(o + RB_D, L), sbn = gvec(o + RB_S, L);
    const bool bscat = o[RB_F] != 0.0;
    if (lflag[l]) {
      Mat G;
      gload(o + RB_G, L, G);
      put(s.tile, L, G);
      put_vec(s.v0, L, Idn);
      wave_sync();
      Idn = vadd(matvec(s.tile, s.v0, L), d);
      wave_sync();
    } else {
      Idn = vadd(vmul(gvec(o + RB_G, L), Idn), d);
    }
    Iup = sbn;
    if (bscat) {
      Mat Rb;
      gload(o + RB_R, L, Rb);
      put(s.tile, L, Rb);
      put_vec(s.v0, L, Idn);
      wave_sync();
      Iup = vadd(matvec(s.tile, s.v0, L), sbn);
      wave_sync();
    }
    emit(l + 1, Iup, Idn);
  }
}

}  // namespace w32
}  // namespace dev
}  // namespace adrt
