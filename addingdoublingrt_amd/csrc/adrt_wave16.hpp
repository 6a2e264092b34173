// adrt_wave16.hpp -- N = 16 building blocks on the matrix cores (gfx950).
//
// One 64-lane wave owns one 16 x 16 problem.  Lane l = j + 16 h (j = column,
// h = 0..3) keeps 4 entries of column j in registers, c[r] = M(row(h, r), j),
// where row() is the output order of the MFMA used for that precision:
//   FP64  v_mfma_f64_16x16x4_f64 : row(h, r) = h + 4 r   (interleaved)
//   FP32  v_mfma_f32_16x16x4_f32 : row(h, r) = 4 h + r   (blocked)
// (both measured on the device).  A product C = X Y takes Y straight from
// registers as the B operand -- register r of lane (j, h) supplies the K index
// row(h, r) -- and X from a padded column-major LDS tile as the A operand, so
// four MFMAs produce C already in the same register layout and products chain
// without any data movement.  Matrices go to HBM in register order
// (lane-major), so every load and store of a record is fully coalesced.
//
// Vectors of length 16 live in LDS slots; element j is also kept in the
// registers of the four lanes of column j.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "adrt_device.hpp"

namespace adrt {
namespace dev {
namespace w16 {

constexpr int N = 16;
constexpr int TS = 17;           // LDS tile column stride (padding against bank conflicts)
constexpr int TILE = N * TS;     // doubles/floats per tile
// every tile access (put, get, get_t, mul, matvec) indexes tile[k * TS + i], 0 <= i, k < N
static_assert((N - 1) * TS + (N - 1) < TILE, "tile index past the tile");

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename R>
struct Mat {  // this lane's 4 entries of a 16 x 16 matrix
  R v[4];
};

template <typename R>
__device__ __forceinline__ int row(int h, int r) {
  if constexpr (sizeof(R) == 8) return h + 4 * r;
  else return 4 * h + r;
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename R>
using vec4 = typename std::conditional<sizeof(R) == 8, d4, f4>::type;

// keeps the scheduler from interleaving neighbouring primitives (bounds the
// number of values in flight, hence the register count and the occupancy)
__device__ __forceinline__ void fence_sched() { __builtin_amdgcn_sched_barrier(0); }

// index of this wave within its workgroup, as a wave-uniform (scalar) value:
// task indices derived from it keep per-task loads and branches scalar
__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

struct Lane {
  int l, j, h;
  __device__ Lane() : l(threadIdx.x & 63), j(threadIdx.x & 15), h((threadIdx.x & 63) >> 4) {}
};

template <typename R>
__device__ __forceinline__ void put(R* tile, const Lane& L, const Mat<R>& m) {
#pragma unroll
  for (int r = 0; r < 4; ++r) tile[L.j * TS + row<R>(L.h, r)] = m.v[r];
}

// this lane's entries of a tile column-major matrix
template <typename R>
__device__ __forceinline__ void get(const R* tile, const Lane& L, Mat<R>& m) {
#pragma unroll
  for (int r = 0; r < 4; ++r) m.v[r] = tile[L.j * TS + row<R>(L.h, r)];
}

// this lane's entries of the transpose of a tile
template <typename R>
__device__ __forceinline__ void get_t(const R* tile, const Lane& L, Mat<R>& m) {
#pragma unroll
  for (int r = 0; r < 4; ++r) m.v[r] = tile[row<R>(L.h, r) * TS + L.j];
}

template <typename R>
__device__ __forceinline__ void gload(const R* g, const Lane& L, Mat<R>& m) {
#pragma unroll
  for (int r = 0; r < 4; ++r) m.v[r] = g[L.l * 4 + r];
}

// the transpose of a matrix stored in register order: element (row(h, r), j) of
// the transpose is element (j, row(h, r)) of the stored matrix, held by lane
// row(h, r) + 16 h' at register r' with row(h', r') = j
template <typename R>
__device__ __forceinline__ void gload_t(const R* g, const Lane& L, Mat<R>& m) {
  int hj, rj;
  if constexpr (sizeof(R) == 8) { hj = L.j & 3; rj = L.j >> 2; }
  else { hj = L.j >> 2; rj = L.j & 3; }
#pragma unroll
  for (int r = 0; r < 4; ++r) m.v[r] = g[(row<R>(L.h, r) + 16 * hj) * 4 + rj];
}

template <typename R>
__device__ __forceinline__ void gstore(R* g, const Lane& L, const Mat<R>& m) {
#pragma unroll
  for (int r = 0; r < 4; ++r) g[L.l * 4 + r] = m.v[r];
}

// C = X Y with X in an LDS tile, Y in registers
template <typename R>
__device__ __forceinline__ void mul(const R* X, const Mat<R>& Y, const Lane& L, Mat<R>& C) {
  fence_sched();
  vec4<R> acc = {R(0), R(0), R(0), R(0)};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const R a = X[row<R>(L.h, r) * TS + L.j];  // A operand: X(i = j-lane index, k = row(h, r))
    acc = mfma(a, Y.v[r], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = acc[r];
  fence_sched();
}

// C = Z + X Y
template <typename R>
__device__ __forceinline__ void mul_add(const R* X, const Mat<R>& Y, const Mat<R>& Z,
                                        const Lane& L, Mat<R>& C) {
  fence_sched();
  vec4<R> acc = {Z.v[0], Z.v[1], Z.v[2], Z.v[3]};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const R a = X[row<R>(L.h, r) * TS + L.j];
    acc = mfma(a, Y.v[r], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = acc[r];
  fence_sched();
}

// ---- pair mode: two problems of order <= 8 in one 16 x 16 problem ----------
// Problem A in rows and columns 0..7, problem B in 8..15, zeros elsewhere
// (block diagonal).  A product of two such matrices needs only the K range of
// each lane's own block: for the FP64 layout, K chunks r = 0, 1 cover rows
// 0..7 and r = 2, 3 rows 8..15, so each lane feeds the two chunks of its block
// (A operand: row j of X at k = row(h, s) + 8 half(j); B operand: register
// s + 2 half(j)) and two instructions give both diagonal blocks.  The other
// output registers (rows of the other block) hold cross products and are
// zeroed by pmask, which keeps the result block diagonal.  Everything else --
// tiles, transposed reads, matrix-vector products, Newton-Schulz, the LU
// solve -- works on the block-diagonal 16 x 16 matrix unchanged (the LU of a
// block-diagonal matrix never mixes the blocks: the multipliers across them
// are exact zeros).
__device__ __forceinline__ int pair_half(const Lane& L) { return L.j >> 3; }

__device__ __forceinline__ void pmask(d4& acc, const Lane& L) {
  const bool hb = pair_half(L) != 0;
  acc[0] = hb ? 0.0 : acc[0];
  acc[1] = hb ? 0.0 : acc[1];
  acc[2] = hb ? acc[2] : 0.0;
  acc[3] = hb ? acc[3] : 0.0;
}
__device__ __forceinline__ void pmask(Mat<double>& m, const Lane& L) {
  const bool hb = pair_half(L) != 0;
  m.v[0] = hb ? 0.0 : m.v[0];
  m.v[1] = hb ? 0.0 : m.v[1];
  m.v[2] = hb ? m.v[2] : 0.0;
  m.v[3] = hb ? m.v[3] : 0.0;
}

// acc += X Y (X in a tile, Y in registers), both block diagonal; cross blocks
// of acc are left unmasked
__device__ __forceinline__ void pmac(d4& acc, const double* X, const Mat<double>& Y,
                                     const Lane& L) {
  fence_sched();
  const int hb = pair_half(L);
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double a = X[(row<double>(L.h, s) + 8 * hb) * TS + L.j];
    const double b = hb ? Y.v[s + 2] : Y.v[s];
    acc = mfma(a, b, acc);
  }
  fence_sched();
}
// acc += X^T Y (X in a tile, read transposed)
__device__ __forceinline__ void pmac_t(d4& acc, const double* X, const Mat<double>& Y,
                                       const Lane& L) {
  fence_sched();
  const int hb = pair_half(L);
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double a = X[L.j * TS + row<double>(L.h, s) + 8 * hb];
    const double b = hb ? Y.v[s + 2] : Y.v[s];
    acc = mfma(a, b, acc);
  }
  fence_sched();
}
__device__ __forceinline__ void pmul(const double* X, const Mat<double>& Y, const Lane& L,
                                     Mat<double>& C) {
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  pmac(acc, X, Y, L);
  pmask(acc, L);
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = acc[r];
}
__device__ __forceinline__ void pmul_add(const double* X, const Mat<double>& Y,
                                         const Mat<double>& Z, const Lane& L, Mat<double>& C) {
  d4 acc = {Z.v[0], Z.v[1], Z.v[2], Z.v[3]};
  pmac(acc, X, Y, L);
  pmask(acc, L);
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = acc[r];
}

// products of one problem per wave (PAIR = false) or of a block-diagonal pair
template <typename R, bool PAIR>
__device__ __forceinline__ void mulx(const R* X, const Mat<R>& Y, const Lane& L, Mat<R>& C) {
  if constexpr (PAIR) pmul(X, Y, L, C);
  else mul<R>(X, Y, L, C);
}
template <typename R, bool PAIR>
__device__ __forceinline__ void mulx_add(const R* X, const Mat<R>& Y, const Mat<R>& Z,
                                         const Lane& L, Mat<R>& C) {
  if constexpr (PAIR) pmul_add(X, Y, Z, L, C);
  else mul_add<R>(X, Y, Z, L, C);
}

// C = I - X Y
template <typename R>
__device__ __forceinline__ void eye_minus_mul(const R* X, const Mat<R>& Y, const Lane& L,
                                              Mat<R>& C) {
  Mat<R> P;
  mul<R>(X, Y, L, P);
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = ((row<R>(L.h, r) == L.j) ? R(1) : R(0)) - P.v[r];
}

template <typename R>
__device__ __forceinline__ R col_reduce(R p, const Lane& L);

// (X v)_j for X in a tile and v in an LDS slot of 16 values: lane group h
// takes k = 4h..4h+3, the four partial sums are added across the column
template <typename R>
__device__ __forceinline__ R matvec(const R* X, const R* v, const Lane& L) {
  fence_sched();
  R s = R(0);
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int k = 4 * L.h + kk;
    s = fma_(X[k * TS + L.j], v[k], s);
  }
  s = col_reduce<R>(s, L);
  fence_sched();
  return s;
}

// (X v_q)_j for NV vectors in LDS slots of 16 values, v + 16 q, split over
// the lane groups as matvec
template <typename R, int NV>
__device__ __forceinline__ void matvecs(const R* X, const R* v, const Lane& L, R (&s)[NV]) {
  fence_sched();
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = R(0);
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int k = 4 * L.h + kk;
    const R x = X[k * TS + L.j];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = fma_(x, v[q * N + k], s[q]);
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = col_reduce<R>(s[q], L);
  fence_sched();
}

// ---- cross-lane helpers (no LDS traffic, no barrier) ----------------------
__device__ __forceinline__ double shfl(double v, int lane) {  // value of `lane`
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_ds_bpermute(lane << 2, (int)b);
  const int hi = __builtin_amdgcn_ds_bpermute(lane << 2, (int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float shfl(float v, int lane) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_ds_bpermute(lane << 2, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double readlane(double v, int lane) {  // lane is wave-uniform
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float readlane(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
// maximum over the 16 lanes of each DPP row (lanes 16h .. 16h+15), in every lane
template <typename R>
__device__ __forceinline__ R row_max(R v) {
  v = __builtin_fmax(v, dpp<0xB1>(v));   // quad_perm [1,0,3,2]
  v = __builtin_fmax(v, dpp<0x4E>(v));   // quad_perm [2,3,0,1]
  v = __builtin_fmax(v, dpp<0x141>(v));  // row_half_mirror
  v = __builtin_fmax(v, dpp<0x140>(v));  // row_mirror
  return v;
}

// sum over the 16 lanes of each DPP row, the same value in every lane of the
// row (pairs, quads, half-row mirror, row mirror: each stage adds two equal-
// order partial sums, so every lane rounds alike)
template <typename R>
__device__ __forceinline__ R row_sum(R v) {
  v = v + dpp<0xB1>(v);
  v = v + dpp<0x4E>(v);
  v = v + dpp<0x141>(v);
  v = v + dpp<0x140>(v);
  return v;
}

// diagonal entry M(j, j) of column j, for all four lanes of the column
template <typename R>
__device__ __forceinline__ R diag(const Mat<R>& m, const Lane& L) {
  // owner of M(j, j) is lane (j, h*) with row(h*, r*) == j
  int oh, orr;
  if constexpr (sizeof(R) == 8) { oh = L.j & 3; orr = L.j >> 2; }
  else { oh = L.j >> 2; orr = L.j & 3; }
  const R sel = orr == 0 ? m.v[0] : orr == 1 ? m.v[1] : orr == 2 ? m.v[2] : m.v[3];
  return shfl(sel, L.j + 16 * oh);
}

// Pivot choice of elimination step k on [A^T | T^T] (row partial pivoting as an LU of
// A^T; the pivot is the first row of largest magnitude among those not yet
// used, and the pivot row is scaled with the corrected reciprocal).  Lane
// (j, h) holds row j of A^T (= column j of A) at columns row(h, .).
template <typename R, int k>
__device__ __forceinline__ int gj_pivot(const Mat<R>& A, const Lane& L, bool used, R& f, R& fp) {
  constexpr int kh = (sizeof(R) == 8) ? (k & 3) : (k >> 2);
  constexpr int kr = (sizeof(R) == 8) ? (k >> 2) : (k & 3);
  const R own = A.v[kr];                           // A(k, j), meaningful in row kh
  const R key = used ? R(-1) : __builtin_fabs(own);
  const R m = row_max<R>(key);
  const unsigned long long hit = __builtin_amdgcn_ballot_w64(key == m);
  const int p = __builtin_ctzll((hit >> (16 * kh)) & 0xFFFFull);
  f = shfl(own, L.j + 16 * kh);                    // multiplier of row j
  fp = readlane(own, p + 16 * kh);                 // pivot
  return p;
}

// X A = T for X, in place on T, as the reference's LU of A^T with partial
// pivoting (src/matrix.hpp:120-124): forward elimination of the rows not yet
// pivoted (pivot rows scaled to a unit diagonal), then back substitution in
// reverse pivot order.  (Eliminating above the pivot as well, Gauss-Jordan,
// gives the same forward error bound but a residual larger by up to
// cond(A); on the near-singular A of the reference's unstable doubling
// starts that cost two digits.)  No LDS: pivots by DPP row maxima and a
// ballot, pivot rows by lane permutes, the final row permutation by one
// permute per register.
template <typename R>
__device__ __forceinline__ void right_solve(Mat<R>& A, Mat<R>& T, const Lane& L) {
  fence_sched();
  bool used = false;
  int src = 0;       // lane column holding row j of the result
  int mystep = N;    // step at which this lane column's row was the pivot
  int piv_of[N];     // pivot row of each step (wave-uniform)
  auto step = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    R f, fp;
    const int p = gj_pivot<R, k>(A, L, used, f, fp);
    piv_of[k] = p;
    const R inv = R(1) / fp;
    const bool piv = (L.j == p);
    const int pl = p + 16 * L.h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const R pa = div_by(shfl(A.v[r], pl), fp, inv);
      const R pt = div_by(shfl(T.v[r], pl), fp, inv);
      A.v[r] = piv ? pa : (used ? A.v[r] : fma_(-f, pa, A.v[r]));
      T.v[r] = piv ? pt : (used ? T.v[r] : fma_(-f, pt, T.v[r]));
    }
    mystep = piv ? k : mystep;
    used = used || piv;
    src = (L.j == k) ? p : src;
  };
  auto back = [&](auto kc) {
    constexpr int k = N - 1 - decltype(kc)::value;  // N-1 .. 1
    constexpr int kh = (sizeof(R) == 8) ? (k & 3) : (k >> 2);
    constexpr int kr = (sizeof(R) == 8) ? (k >> 2) : (k & 3);
    const R f = shfl(A.v[kr], L.j + 16 * kh);  // U(row j, k)
    const int pl = piv_of[k] + 16 * L.h;
    const bool upd = mystep < k;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const R pt = shfl(T.v[r], pl);
      T.v[r] = upd ? fma_(-f, pt, T.v[r]) : T.v[r];
    }
  };
  [&]<int... K>(std::integer_sequence<int, K...>) {
    (step(std::integral_constant<int, K>{}), ...);
  }(std::make_integer_sequence<int, N>{});
  [&]<int... K>(std::integer_sequence<int, K...>) {
    (back(std::integral_constant<int, K>{}), ...);
  }(std::make_integer_sequence<int, N - 1>{});
  // row mystep of the reduced system is row mystep of X^T: lane (j, h) takes
  // column j of X from the lane column pivoted at step j
  const int sl = src + 16 * L.h;
#pragma unroll
  for (int r = 0; r < 4; ++r) T.v[r] = shfl(T.v[r], sl);
  fence_sched();
}

// the largest over the wave of |x| for the FP32 bit patterns u = bits(|x|) of
// each lane, in every lane: the bit patterns of non-negative floats order as
// unsigned integers as the values do (a NaN, above +inf, wins)
__device__ __forceinline__ float wave_max_abs_bits(unsigned u) {
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0xB1, 0xF, 0xF, false));
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x4E, 0xF, 0xF, false));
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x141, 0xF, 0xF, false));
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x140, 0xF, 0xF, false));
  const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)u, 0);
  const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)u, 16);
  const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)u, 32);
  const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)u, 48);
  return __builtin_bit_cast(float, max(max(a, b), max(c, d)));
}
__device__ __forceinline__ unsigned abs_bits(float x) {
  return __builtin_bit_cast(unsigned, __builtin_fabsf(x));
}

// N max |E_ij| over the wave (bounds the infinity norm of E), wave-uniform
template <typename R>
__device__ __forceinline__ float residual_size(const Mat<R>& E) {
  unsigned m = 0u;
#pragma unroll
  for (int r = 0; r < 4; ++r) m = max(m, abs_bits((float)E.v[r]));
  return wave_max_abs_bits(m) * (float)N;
}

// Residual sizes at which the Newton-Schulz inverse stops: the iterate is
// taken as is, after one Newton step (residual e^2) or after one third-order
// step (residual e^3), each leaving a residual below FP64 rounding.
constexpr float NS_DONE = 1e-16f, NS_NEWTON = 1e-8f, NS_CUBIC = 1e-5f;
constexpr int NS_MAX_STEPS = 8;

// X = (I - P)^-1 for P in registers and in the tile tP (tX: scratch tile).
// Newton-Schulz iteration on matrix cores from the Neumann start X0 = I + P,
// whose residual E0 = I - (I - P) X0 is P^2:
//   X <- X + X E          (new residual E^2)
//   X <- X + X E + X E E  (new residual E^3, once e = N max|E| <= NS_CUBIC)
// with E = I - X + P X recomputed from each iterate, so rounding does not
// accumulate.  The doubling's thin layers give e ~ (tau0 |G|)^4 and stop at
// once; where the start is poor (e > 1/2: (I - P) close to singular) or the
// iteration has not finished in NS_MAX_STEPS steps, LU with partial
// pivoting (right_solve) solves X (I - P) = I instead.  Every branch is
// wave-uniform.  PAIR: P is a block-diagonal pair (pair mode above); the
// residual test then covers both blocks.  Returns true if the LU solve ran.
// With FALLBACK = false the fallback is left to the caller: the function returns true at once, X unset,
// so that a caller with a right-hand side can solve for it directly (a solve
// keeps the error of an ill-conditioned I - P at that of an LU solve; the
// product with a computed inverse loses up to cond(I - P) more).
template <typename R, bool FALLBACK = true, bool PAIR = false>
__device__ __forceinline__ bool inverse_ns(const Mat<R>& P, const R* tP, R* tX, Mat<R>& X,
                                           const Lane& L) {
  Mat<R> E;
  mulx<R, PAIR>(tP, P, L, E);  // P^2
#pragma unroll
  for (int r = 0; r < 4; ++r) X.v[r] = ((row<R>(L.h, r) == L.j) ? R(1) : R(0)) + P.v[r];
  float e = residual_size<R>(E);
  for (int it = 0; e > NS_DONE; ++it) {
    if (!(e <= 0.5f) || it == NS_MAX_STEPS) {
      if constexpr (!FALLBACK) return true;
      // P again from its tile, so no register holds it through the iteration
      Mat<R> A;
      get<R>(tP, L, A);
      wave_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const R d = (row<R>(L.h, r) == L.j) ? R(1) : R(0);
        A.v[r] = d - A.v[r];
        X.v[r] = d;
      }
      right_solve<R>(A, X, L);
      return true;
    }
    put<R>(tX, L, X);
    wave_sync();
    if (e <= NS_NEWTON) {
      mulx_add<R, PAIR>(tX, E, X, L, X);
      wave_sync();
      break;
    }
    Mat<R> XE;
    mulx<R, PAIR>(tX, E, L, XE);
    wave_sync();
    if (e <= NS_CUBIC) {
      put<R>(tX, L, XE);
      wave_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) X.v[r] += XE.v[r];
      mulx_add<R, PAIR>(tX, E, X, L, X);
      wave_sync();
      break;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) X.v[r] += XE.v[r];
    Mat<R> PX;
    mulx<R, PAIR>(tP, X, L, PX);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      E.v[r] = (((row<R>(L.h, r) == L.j) ? R(1) : R(0)) - X.v[r]) + PX.v[r];
    e = residual_size<R>(E);
  }
  wave_sync();
  return false;
}

// D = T (I - P)^-1 for P in the tile tP, by LU with partial pivoting
// on D (I - P) = T (the fallback of inverse_ns<R, false>).
template <typename R>
__device__ __forceinline__ void right_solve_tile(const R* tP, const Mat<R>& T, Mat<R>& D,
                                                 const Lane& L) {
  Mat<R> A;
  get<R>(tP, L, A);
  wave_sync();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    A.v[r] = ((row<R>(L.h, r) == L.j) ? R(1) : R(0)) - A.v[r];
    D.v[r] = T.v[r];
  }
  right_solve<R>(A, D, L);
}

// G = (I - P)^-1 T for P in the tile tP, by LU with partial pivoting on the
// transposed system G^T (I - P)^T = T^T (tS: scratch tile)
template <typename R>
__device__ __forceinline__ void left_solve_mat_tile(const R* tP, const Mat<R>& T, Mat<R>& G,
                                                    R* tS, const Lane& L) {
  Mat<R> A;
  get_t<R>(tP, L, A);
  put<R>(tS, L, T);
  wave_sync();
  get_t<R>(tS, L, G);
  wave_sync();
#pragma unroll
  for (int r = 0; r < 4; ++r) A.v[r] = ((row<R>(L.h, r) == L.j) ? R(1) : R(0)) - A.v[r];
  right_solve<R>(A, G, L);
  put<R>(tS, L, G);
  wave_sync();
  get_t<R>(tS, L, G);
  wave_sync();
}

// Q = P (I - P)^-1 = (I - P)^-1 - I, for P in Q on entry and in the tile tP
// (tX: scratch tile), in Q form so that a small Q keeps its relative
// precision in FP32: the Newton-Schulz iteration of X = I + Q from Q0 = P,
//   E = P - Q + P Q  (= I - (I - P) X),   Q <- Q + E + Q E  (residual E^2),
// until N max|E| <= 1e-5 N max|P| -- a relative accuracy that one refinement
// step with an FP64 residual then carries to FP32 rounding.  E0 = P^2; an
// E0 past 1/2 (I - P close to singular) or 8 steps fall back to the LU solve.
template <typename R>
__device__ __forceinline__ void q_form_ns(Mat<R>& Q, const R* tP, R* tX, const Lane& L) {
  const float pmax = residual_size<R>(Q);
  Mat<R> E;
  mul<R>(tP, Q, L, E);  // P^2
  float e = residual_size<R>(E);
  for (int it = 0; e > 1e-5f * pmax; ++it) {
    if (!(e <= 0.5f) || it == NS_MAX_STEPS) {
      Mat<R> A;
      get<R>(tP, L, A);
      wave_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Q.v[r] = A.v[r];
        A.v[r] = ((row<R>(L.h, r) == L.j) ? R(1) : R(0)) - A.v[r];
      }
      right_solve<R>(A, Q, L);
      return;
    }
    put<R>(tX, L, Q);
    wave_sync();
    {
      Mat<R> Z;
#pragma unroll
      for (int r = 0; r < 4; ++r) Z.v[r] = Q.v[r] + E.v[r];
      mul_add<R>(tX, E, Z, L, Q);
    }
    wave_sync();
    Mat<R> PQ, Pm;
    mul<R>(tP, Q, L, PQ);
    get<R>(tP, L, Pm);
#pragma unroll
    for (int r = 0; r < 4; ++r) E.v[r] = (Pm.v[r] - Q.v[r]) + PQ.v[r];
    e = residual_size<R>(E);
  }
  wave_sync();
}

// x = M^-1 b for A = M^T in registers (destroyed), b element j (all lanes of
// column j); returns x_j.  x^T M^T = b^T is a right solve with A = M^T, by
// the forward elimination and back substitution of right_solve.
template <typename R>
__device__ __forceinline__ R left_solve_t(Mat<R>& A, R b, const Lane& L) {
  bool used = false;
  int src = 0;
  int mystep = N;
  int piv_of[N];
  R t0 = b;
  auto step = [&](auto kc) {
    constexpr int k = decltype(kc)::value;
    R f, fp;
    const int p = gj_pivot<R, k>(A, L, used, f, fp);
    piv_of[k] = p;
    const R inv = R(1) / fp;
    const bool piv = (L.j == p);
    const int pl = p + 16 * L.h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const R pa = div_by(shfl(A.v[r], pl), fp, inv);
      A.v[r] = piv ? pa : (used ? A.v[r] : fma_(-f, pa, A.v[r]));
    }
    const R ptv = div_by(readlane(t0, p), fp, inv);
    t0 = piv ? ptv : (used ? t0 : fma_(-f, ptv, t0));
    mystep = piv ? k : mystep;
    used = used || piv;
    src = (L.j == k) ? p : src;
  };
  auto back = [&](auto kc) {
    constexpr int k = N - 1 - decltype(kc)::value;  // N-1 .. 1
    constexpr int kh = (sizeof(R) == 8) ? (k & 3) : (k >> 2);
    constexpr int kr = (sizeof(R) == 8) ? (k >> 2) : (k & 3);
    const R f = shfl(A.v[kr], L.j + 16 * kh);
    const R ptv = readlane(t0, piv_of[k]);
    t0 = (mystep < k) ? fma_(-f, ptv, t0) : t0;
  };
  [&]<int... K>(std::integer_sequence<int, K...>) {
    (step(std::integral_constant<int, K>{}), ...);
  }(std::make_integer_sequence<int, N>{});
  [&]<int... K>(std::integer_sequence<int, K...>) {
    (back(std::integral_constant<int, K>{}), ...);
  }(std::make_integer_sequence<int, N - 1>{});
  return shfl(t0, src + 16 * L.h);
}

// x = M^-1 b: M by registers (lane j needs column j of M^T, so M goes through
// the tile)
template <typename R>
__device__ __forceinline__ R left_solve(const Mat<R>& M, R b, const Lane& L, R* tile) {
  put<R>(tile, L, M);
  wave_sync();
  Mat<R> A;
  get_t<R>(tile, L, A);
  wave_sync();
  return left_solve_t<R>(A, b, L);
}

// x = (I - P)^-1 b for P in the tile tP, by LU (the fallback of
// inverse_ns<R, false> for one right-hand side)
template <typename R>
__device__ __forceinline__ R left_solve_tile(const R* tP, R b, const Lane& L) {
  Mat<R> A;
  get_t<R>(tP, L, A);
  wave_sync();
#pragma unroll
  for (int r = 0; r < 4; ++r) A.v[r] = ((row<R>(L.h, r) == L.j) ? R(1) : R(0)) - A.v[r];
  return left_solve_t<R>(A, b, L);
}

// ---- reverse-pass building blocks -------------------------------------------

// sum over the four lanes of column j (every lane of the column gets the same
// value: the additions pair (h0 + h1) + (h2 + h3) in every lane)
template <typename R>
__device__ __forceinline__ R col_reduce(R p, const Lane& L) {
  p += shfl(p, L.l ^ 16);
  p += shfl(p, L.l ^ 32);
  return p;
}

// sum over all 64 lanes (fixed order, identical in every lane): the DPP row
// sum of each 16-lane row, then the four rows across the column
template <typename R>
__device__ __forceinline__ R wave_sum(R p, const Lane& L) {
  return col_reduce<R>(row_sum<R>(p), L);
}

// (M^T v)_j with M in registers and v in an LDS slot of 16 values
template <typename R>
__device__ __forceinline__ R tmatvec(const Mat<R>& M, const R* v, const Lane& L) {
  R p = R(0);
#pragma unroll
  for (int r = 0; r < 4; ++r) p = fma_(M.v[r], v[row<R>(L.h, r)], p);
  return col_reduce<R>(p, L);
}

// acc += X Y (X in an LDS tile, Y in registers)
template <typename R>
__device__ __forceinline__ void mac(vec4<R>& acc, const R* X, const Mat<R>& Y, const Lane& L) {
  fence_sched();
#pragma unroll
  for (int r = 0; r < 4; ++r) acc = mfma(X[row<R>(L.h, r) * TS + L.j], Y.v[r], acc);
  fence_sched();
}

// acc += X^T Y (X in an LDS tile, read transposed)
template <typename R>
__device__ __forceinline__ void mac_t(vec4<R>& acc, const R* X, const Mat<R>& Y, const Lane& L) {
  fence_sched();
#pragma unroll
  for (int r = 0; r < 4; ++r) acc = mfma(X[L.j * TS + row<R>(L.h, r)], Y.v[r], acc);
  fence_sched();
}

// acc += sum_k a_k b_k^T for four vectors given element-wise (element j in
// every lane of column j): one K = 4 matrix instruction, lane (j, h) feeding
// a_h[j] and b_h[j]
template <typename R>
__device__ __forceinline__ void outer4(vec4<R>& acc, R a0, R a1, R a2, R a3, R b0, R b1, R b2,
                                       R b3, const Lane& L) {
  const R av = L.h == 0 ? a0 : L.h == 1 ? a1 : L.h == 2 ? a2 : a3;
  const R bv = L.h == 0 ? b0 : L.h == 1 ? b1 : L.h == 2 ? b2 : b3;
  acc = mfma(av, bv, acc);
}

template <typename R>
__device__ __forceinline__ vec4<R> acc_of(const Mat<R>& m) {
  return vec4<R>{m.v[0], m.v[1], m.v[2], m.v[3]};
}
template <typename R>
__device__ __forceinline__ Mat<R> mat_of(const vec4<R>& a) {
  return Mat<R>{{a[0], a[1], a[2], a[3]}};
}
template <typename R>
__device__ __forceinline__ vec4<R> zero4() {
  return vec4<R>{R(0), R(0), R(0), R(0)};
}

// ordered sum over the 16 columns (one value per column), as the reference's
// quadrature loops
template <typename R>
__device__ __forceinline__ R colsum(R v, const Lane& L, R* slot) {
  if (L.h == 0) slot[L.j] = v;
  wave_sync();
  R s = R(0);
#pragma unroll 4
  for (int i = 0; i < N; ++i) s += slot[i];
  wave_sync();
  return s;
}

}  // namespace w16
}  // namespace dev
}  // namespace adrt
