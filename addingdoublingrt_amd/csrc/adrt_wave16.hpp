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

#include "adrt_device.hpp"

namespace adrt {
namespace dev {
namespace w16 {

constexpr int N = 16;
constexpr int TS = 17;           // LDS tile column stride (padding against bank conflicts)
constexpr int TILE = N * TS;     // doubles/floats per tile

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

template <typename R>
__device__ __forceinline__ void gstore(R* g, const Lane& L, const Mat<R>& m) {
#pragma unroll
  for (int r = 0; r < 4; ++r) g[L.l * 4 + r] = m.v[r];
}

// C = X Y with X in an LDS tile, Y in registers
template <typename R>
__device__ __forceinline__ void mul(const R* X, const Mat<R>& Y, const Lane& L, Mat<R>& C) {
  vec4<R> acc = {R(0), R(0), R(0), R(0)};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const R a = X[row<R>(L.h, r) * TS + L.j];  // A operand: X(i = j-lane index, k = row(h, r))
    acc = mfma(a, Y.v[r], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = acc[r];
}

// C = Z + X Y
template <typename R>
__device__ __forceinline__ void mul_add(const R* X, const Mat<R>& Y, const Mat<R>& Z,
                                        const Lane& L, Mat<R>& C) {
  vec4<R> acc = {Z.v[0], Z.v[1], Z.v[2], Z.v[3]};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const R a = X[row<R>(L.h, r) * TS + L.j];
    acc = mfma(a, Y.v[r], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) C.v[r] = acc[r];
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

// (X v)_j for X in a tile and v in an LDS slot of 16 values
template <typename R>
__device__ __forceinline__ R matvec(const R* X, const R* v, const Lane& L) {
  R s = R(0);
#pragma unroll
  for (int k = 0; k < N; ++k) s = fma_(X[k * TS + L.j], v[k], s);
  return s;
}

template <typename R, int NV>
__device__ __forceinline__ void matvecs(const R* X, const R* v, const Lane& L, R (&s)[NV]) {
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = R(0);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const R x = X[k * TS + L.j];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = fma_(x, v[q * N + k], s[q]);
  }
}

// diagonal entry M(j, j) of column j, for all four lanes of the column
template <typename R>
__device__ __forceinline__ R diag(const Mat<R>& m, const Lane& L, R* slot) {
  R d = R(0);
  bool own = false;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const bool hit = row<R>(L.h, r) == L.j;
    d = hit ? m.v[r] : d;
    own = own || hit;
  }
  if (own) slot[L.j] = d;
  wave_sync();
  const R out = slot[L.j];
  wave_sync();
  return out;
}

// X A = T for X, in place on T (Gauss-Jordan on [A^T | T^T], row partial
// pivoting as an LU of A^T, corrected-reciprocal scaling of the pivot row).
// Lane (j, h) holds its entries of column j of A (= row j of A^T) and of T.
// The pivot index is wave-uniform (one problem per wave).
// LDS: fv (16), pv (16), prow (2 x 64 in lane order), perm tile (TILE).
template <typename R>
__device__ __forceinline__ void right_solve(Mat<R>& A, Mat<R>& T, const Lane& L, R* fv, R* pv,
                                            R* prow, R* perm) {
  bool used = false;
  int mystep = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    // owner of A(k, j): row(h, r) == k
    int kh, kr;
    if constexpr (sizeof(R) == 8) { kh = k % 4; kr = k / 4; }
    else { kh = k / 4; kr = k % 4; }
    if (L.h == kh) {
      const R f = A.v[kr];
      fv[L.j] = f;
      pv[L.j] = used ? R(-1) : __builtin_fabs(f);
    }
    wave_sync();
    const R f = fv[L.j];
    int p = 0;
    R best = pv[0];
#pragma unroll
    for (int i = 1; i < N; ++i) {
      const R v = pv[i];
      const bool gt = v > best;
      best = gt ? v : best;
      p = gt ? i : p;
    }
    p = __builtin_amdgcn_readfirstlane(p);
    const bool piv = (L.j == p);
    if (piv) {  // the four lanes of the pivot column scale and publish the pivot row
      const R inv = R(1) / f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        A.v[r] = div_by(A.v[r], f, inv);
        T.v[r] = div_by(T.v[r], f, inv);
        prow[L.h * 4 + r] = A.v[r];
        prow[64 + L.h * 4 + r] = T.v[r];
      }
      used = true;
      mystep = k;
    }
    wave_sync();
    if (!piv) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        A.v[r] = fma_(-f, prow[L.h * 4 + r], A.v[r]);
        T.v[r] = fma_(-f, prow[64 + L.h * 4 + r], T.v[r]);
      }
    }
    wave_sync();
  }
  // lane (j, h) holds rows row(h,.) of column `mystep` of X
#pragma unroll
  for (int r = 0; r < 4; ++r) perm[mystep * TS + row<R>(L.h, r)] = T.v[r];
  wave_sync();
#pragma unroll
  for (int r = 0; r < 4; ++r) T.v[r] = perm[L.j * TS + row<R>(L.h, r)];
  wave_sync();
}

// x = M^-1 b: M by registers, b element j; returns x_j (all lanes of column j)
template <typename R>
__device__ __forceinline__ R left_solve(const Mat<R>& M, R b, const Lane& L, R* fv, R* pv,
                                        R* prow, R* tile) {
  // x^T M^T = b^T: right solve with A = M^T (lane j needs column j of M^T)
  put<R>(tile, L, M);
  wave_sync();
  Mat<R> A;
  get_t<R>(tile, L, A);
  wave_sync();
  bool used = false;
  int mystep = 0;
  R t0 = b;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int kh, kr;
    if constexpr (sizeof(R) == 8) { kh = k % 4; kr = k / 4; }
    else { kh = k / 4; kr = k % 4; }
    if (L.h == kh) {
      const R f = A.v[kr];
      fv[L.j] = f;
      pv[L.j] = used ? R(-1) : __builtin_fabs(f);
    }
    wave_sync();
    const R f = fv[L.j];
    int p = 0;
    R best = pv[0];
#pragma unroll
    for (int i = 1; i < N; ++i) {
      const R v = pv[i];
      const bool gt = v > best;
      best = gt ? v : best;
      p = gt ? i : p;
    }
    p = __builtin_amdgcn_readfirstlane(p);
    const bool piv = (L.j == p);
    if (piv) {
      const R inv = R(1) / f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        A.v[r] = div_by(A.v[r], f, inv);
        prow[L.h * 4 + r] = A.v[r];
      }
      t0 = div_by(t0, f, inv);
      if (L.h == 0) prow[64] = t0;
      used = true;
      mystep = k;
    }
    wave_sync();
    if (!piv) {
#pragma unroll
      for (int r = 0; r < 4; ++r) A.v[r] = fma_(-f, prow[L.h * 4 + r], A.v[r]);
      t0 = fma_(-f, prow[64], t0);
    }
    wave_sync();
  }
  if (L.h == 0) fv[mystep] = t0;
  wave_sync();
  const R x = fv[L.j];
  wave_sync();
  return x;
}

// ordered sum over the 16 columns (one value per column), as the reference's
// quadrature loops
template <typename R>
__device__ __forceinline__ R colsum(R v, const Lane& L, R* slot) {
  if (L.h == 0) slot[L.j] = v;
  wave_sync();
  R s = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) s += slot[i];
  wave_sync();
  return s;
}

}  // namespace w16
}  // namespace dev
}  // namespace adrt
