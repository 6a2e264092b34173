// adrt_device.hpp -- register/LDS building blocks for the gfx950 kernels.
//
// Task-group layout: a group of G = N*H consecutive lanes of one wave64 owns
// one N x N problem (one wavenumber-layer, or one wavenumber).  Lane (j, h)
// (j = column, h = 0..H-1) keeps rows h*RP .. h*RP+RP-1 (RP = N/H) of column j
// of every matrix in registers: c[r] = M(h*RP + r, j).  H is 1 for N <= 8 and 2
// for N >= 16, which keeps the live state of the doubling / adding recursions
// under the 256-VGPR budget; the two lanes of a column are neighbours, so
// they exchange halves with one DPP quad permutation.
// Vectors are distributed by column: both lanes of column j hold element j.
//
// Cross-lane traffic goes through a per-group LDS window.  A group never spans
// two waves, so ordering LDS writes before reads of other lanes of the group
// only needs a wavefront-scope fence (LDS operations of one wave complete in
// order); no s_barrier is used anywhere.
#pragma once

#include <hip/hip_runtime.h>

namespace adrt {
namespace dev {

constexpr double kPi = 3.14159265358979323846;

template <int N>
struct Geo {
  static constexpr int H = N >= 8 ? 2 : 1;  // lanes per column
  static constexpr int RP = N / H;           // rows per lane
  static constexpr int G = N * H;            // lanes per group
  static constexpr int GPB = 256 / G;        // groups per 256-thread block
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R>
__device__ __forceinline__ R fma_(R a, R b, R c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// Value held by the neighbour lane (lane ^ 1): DPP quad_perm [1,0,3,2].
__device__ __forceinline__ int swap_pair_i(int x) {
  return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ double swap_pair(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = swap_pair_i((int)(b & 0xffffffffLL));
  const int hi = swap_pair_i((int)(b >> 32));
  const long long r = ((long long)(unsigned)lo) | ((long long)hi << 32);
  return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ float swap_pair(float x) {
  return __builtin_bit_cast(float, swap_pair_i(__builtin_bit_cast(int, x)));
}

// ---- column slices ----------------------------------------------------------

// Store this lane's rows of column j into a column-major N x N tile.
template <int N, typename R>
__device__ __forceinline__ void put_col(R* tile, int j, int h, const R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
  R* p = tile + j * N + h * RP;
#pragma unroll
  for (int r = 0; r < RP; ++r) p[r] = c[r];
}

template <int N, typename R>
__device__ __forceinline__ void load_col(const R* m, int j, int h, R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
  const R* p = m + j * N + h * RP;
#pragma unroll
  for (int r = 0; r < RP; ++r) c[r] = p[r];
}

template <int N, typename R>
__device__ __forceinline__ void store_col(R* m, int j, int h, const R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
  R* p = m + j * N + h * RP;
#pragma unroll
  for (int r = 0; r < RP; ++r) p[r] = c[r];
}

// Rows h*RP.. of row j of a column-major tile (= this lane's slice of column j
// of the transpose).
template <int N, typename R>
__device__ __forceinline__ void get_row(const R* tile, int j, int h, R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
#pragma unroll
  for (int r = 0; r < RP; ++r) c[r] = tile[(h * RP + r) * N + j];
}

// Whole column j (all N rows) in this lane: own slice plus the neighbour's.
template <int N, typename R>
__device__ __forceinline__ void full_col(const R (&c)[Geo<N>::RP], int h, R (&y)[N]) {
  constexpr int H = Geo<N>::H, RP = Geo<N>::RP;
  if constexpr (H == 1) {
#pragma unroll
    for (int r = 0; r < N; ++r) y[r] = c[r];
  } else {
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const R other = swap_pair(c[r]);
      y[r] = h == 0 ? c[r] : other;
      y[RP + r] = h == 0 ? other : c[r];
    }
  }
}

// ---- products ---------------------------------------------------------------

// c = X * y : X is an LDS tile (column-major), y this lane's slice of column j
// of Y.  The k loop is software-pipelined by hand (column k+1 of X is read
// while column k is consumed) and fenced with scheduling barriers; left alone
// the scheduler hoists all N*N loads and spills.
template <int N, typename R>
__device__ __forceinline__ void mul(const R* X, const R (&y)[Geo<N>::RP], int h,
                                    R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
  R yf[N];
  full_col<N, R>(y, h, yf);
  const R* Xh = X + h * RP;
  R cur[RP], nxt[RP];
#pragma unroll
  for (int r = 0; r < RP; ++r) { c[r] = R(0); cur[r] = Xh[r]; }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k + 1 < N) {
#pragma unroll
      for (int r = 0; r < RP; ++r) nxt[r] = Xh[(k + 1) * N + r];
    }
    __builtin_amdgcn_sched_barrier(0);
    const R yk = yf[k];
#pragma unroll
    for (int r = 0; r < RP; ++r) c[r] = fma_(cur[r], yk, c[r]);
    __builtin_amdgcn_sched_barrier(0);
    if (k + 1 < N) {
#pragma unroll
      for (int r = 0; r < RP; ++r) cur[r] = nxt[r];
    }
  }
}

// c += X * y
template <int N, typename R>
__device__ __forceinline__ void mul_acc(const R* X, const R (&y)[Geo<N>::RP], int h,
                                        R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
  R t[RP];
  mul<N, R>(X, y, h, t);
#pragma unroll
  for (int r = 0; r < RP; ++r) c[r] = c[r] + t[r];
}

// c = I(:,j) - X * y
template <int N, typename R>
__device__ __forceinline__ void eye_minus_mul(const R* X, const R (&y)[Geo<N>::RP], int j,
                                              int h, R (&c)[Geo<N>::RP]) {
  constexpr int RP = Geo<N>::RP;
  R t[RP];
  mul<N, R>(X, y, h, t);
#pragma unroll
  for (int r = 0; r < RP; ++r) c[r] = ((h * RP + r == j) ? R(1) : R(0)) - t[r];
}

// Distributed mat-vecs with the matrix in an LDS tile and NV vectors in
// consecutive LDS slots of N values (v + q*N): s[q] = (X v_q)_i for row i, in
// both lanes of column i.  With two lanes per column each sums half of the k
// range and the halves are added across the pair (the two lanes add the same
// two numbers).  Pipelined like mul() so that the loads of several mat-vecs
// are not hoisted together.
template <int N, int NV, typename R>
__device__ __forceinline__ void matvecs(const R* X, const R* v, int i, R (&s)[NV]) {
  constexpr int H = Geo<N>::H, KN = N / H;
  const int k0 = H == 1 ? 0 : (int)(threadIdx.x % H) * KN;
  R x = X[k0 * N + i], vv[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) { s[q] = R(0); vv[q] = v[q * N + k0]; }
#pragma unroll
  for (int kk = 0; kk < KN; ++kk) {
    R xn = R(0), vn[NV];
    if (kk + 1 < KN) {
      xn = X[(k0 + kk + 1) * N + i];
#pragma unroll
      for (int q = 0; q < NV; ++q) vn[q] = v[q * N + k0 + kk + 1];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = fma_(x, vv[q], s[q]);
    __builtin_amdgcn_sched_barrier(0);
    if (kk + 1 < KN) {
      x = xn;
#pragma unroll
      for (int q = 0; q < NV; ++q) vv[q] = vn[q];
    }
  }
  if constexpr (H == 2) {
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = s[q] + swap_pair(s[q]);
  }
}

template <int N, typename R>
__device__ __forceinline__ R matvec(const R* X, const R* v, int i) {
  R s[1];
  matvecs<N, 1, R>(X, v, i, s);
  return s[0];
}

// x / f from a precomputed 1/f plus one residual correction: as accurate as a
// true division (the residual x - y f is exact with an FMA) at 3 FMAs instead
// of a full IEEE division sequence.  Scaling the pivot row by the bare
// reciprocal instead loses up to ~10x accuracy on ill-conditioned (I - R R)
// systems against the reference's LU, which divides.
template <typename R>
__device__ __forceinline__ R div_by(R x, R f, R inv) {
  const R y = x * inv;
  return fma_(fma_(-y, f, x), inv, y);
}

// ---- right solve ------------------------------------------------------------
//
// X * A = T for X (the reference's rightSolveMatrix, src/matrix.hpp:120-124,
// which factors A^T with row partial pivoting).  Lane (j, h) holds its slice of
// column j of A (= row j of A^T) in a[] and of column j of T (= row j of T^T)
// in t[].  LU of [A^T | T^T] with partial pivoting over the rows not used yet:
// at step k the pivot is the first largest |A^T(i,k)|, the pivot row is scaled
// to a unit diagonal and eliminated from the rows not yet used; back
// substitution in reverse pivot order then clears the rows pivoted earlier.
// With LU = false the pivot row is eliminated from every other row at once
// (Gauss-Jordan): the same forward error bound, a residual larger by up to
// cond(A), no back-substitution pass; the doubling uses it, where it runs
// every step and A = I - R^2 is far from singular outside the reference's
// unstable starts.
// Every row's multiplier is its own A^T(i,k) (static register index, fetched
// from the neighbour lane when it sits in the other half), and the pivot row
// travels through LDS.  On return t[] holds this lane's slice of column j of X.
//
// LDS: prow (2N values), pval (N values), perm (N*N values).
template <int N, typename R>
__device__ __forceinline__ R column_entry(const R (&a)[Geo<N>::RP], int k, int h) {
  constexpr int H = Geo<N>::H, RP = Geo<N>::RP;
  const int kh = k / RP, kr = k % RP;
  R f = a[kr];
  if constexpr (H == 2) {
    const R other = swap_pair(a[kr]);
    f = (h == kh) ? a[kr] : other;
  }
  return f;
}

template <int N, typename R, bool LU = true>
__device__ __forceinline__ void right_solve(R (&a)[Geo<N>::RP], R (&t)[Geo<N>::RP], int j,
                                            int h, R* prow, R* pval, R* perm) {
  constexpr int RP = Geo<N>::RP;
  bool used = false;
  int mystep = 0;
  int piv_of[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const R f = column_entry<N, R>(a, k, h);
    if (h == 0) pval[j] = used ? R(-1) : __builtin_fabs(f);
    wave_sync();
    int p = 0;
    R best = pval[0];
#pragma unroll
    for (int i = 1; i < N; ++i) {
      const R v = pval[i];
      if (v > best) { best = v; p = i; }
    }
    piv_of[k] = p;
    const bool was_used = used;
    if (j == p) {
      const R inv = R(1) / f;
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        a[r] = div_by(a[r], f, inv);
        prow[h * RP + r] = a[r];
        t[r] = div_by(t[r], f, inv);
        prow[N + h * RP + r] = t[r];
      }
      used = true;
      mystep = k;
    }
    wave_sync();
    if (j != p && (!LU || !was_used)) {
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        a[r] = fma_(-f, prow[h * RP + r], a[r]);
        t[r] = fma_(-f, prow[N + h * RP + r], t[r]);
      }
    }
    wave_sync();
  }
  // back substitution: the row pivoted at step k is final once the later
  // steps are done; it is then cleared from the rows pivoted before it
#pragma unroll
  for (int k = LU ? N - 1 : 0; k > 0; --k) {
    const R f = column_entry<N, R>(a, k, h);
    if (j == piv_of[k]) {
#pragma unroll
      for (int r = 0; r < RP; ++r) prow[N + h * RP + r] = t[r];
    }
    wave_sync();
    if (mystep < k) {
#pragma unroll
      for (int r = 0; r < RP; ++r) t[r] = fma_(-f, prow[N + h * RP + r], t[r]);
    }
    wave_sync();
  }
  // row j of the reduced system is X's column `mystep`
  R* dst = perm + mystep * N + h * RP;
#pragma unroll
  for (int r = 0; r < RP; ++r) dst[r] = t[r];
  wave_sync();
  const R* src = perm + j * N + h * RP;
#pragma unroll
  for (int r = 0; r < RP; ++r) t[r] = src[r];
  wave_sync();
}

// Same with a single right-hand-side row: t0 is T(0, j) (replicated in both
// lanes of column j); returns X(0, j).
template <int N, typename R>
__device__ __forceinline__ R right_solve_vec(R (&a)[Geo<N>::RP], R t0, int j, int h, R* prow,
                                             R* pval, R* perm) {
  constexpr int RP = Geo<N>::RP;
  bool used = false;
  int mystep = 0;
  int piv_of[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const R f = column_entry<N, R>(a, k, h);
    if (h == 0) pval[j] = used ? R(-1) : __builtin_fabs(f);
    wave_sync();
    int p = 0;
    R best = pval[0];
#pragma unroll
    for (int i = 1; i < N; ++i) {
      const R v = pval[i];
      if (v > best) { best = v; p = i; }
    }
    piv_of[k] = p;
    const bool was_used = used;
    if (j == p) {
      const R inv = R(1) / f;
#pragma unroll
      for (int r = 0; r < RP; ++r) {
        a[r] = div_by(a[r], f, inv);
        prow[h * RP + r] = a[r];
      }
      t0 = div_by(t0, f, inv);
      if (h == 0) prow[N] = t0;
      used = true;
      mystep = k;
    }
    wave_sync();
    if (j != p && !was_used) {
#pragma unroll
      for (int r = 0; r < RP; ++r) a[r] = fma_(-f, prow[h * RP + r], a[r]);
      t0 = fma_(-f, prow[N], t0);
    }
    wave_sync();
  }
#pragma unroll
  for (int k = N - 1; k > 0; --k) {
    const R f = column_entry<N, R>(a, k, h);
    if (j == piv_of[k] && h == 0) prow[N] = t0;
    wave_sync();
    if (mystep < k) t0 = fma_(-f, prow[N], t0);
    wave_sync();
  }
  if (h == 0) perm[mystep] = t0;
  wave_sync();
  const R x = perm[j];
  wave_sync();
  return x;
}

// Sum of one value per column over the group, in column order 0..N-1 (the
// order the reference accumulates its quadrature sums in).  Every lane gets it.
template <int N, typename R>
__device__ __forceinline__ R group_sum_ordered(R v, int j, int h, R* slot) {
  if (h == 0) slot[j] = v;
  wave_sync();
  R s = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) s += slot[i];
  wave_sync();
  return s;
}

// Diagonal element M(j,j) of column j without dynamic register indexing.
template <int N, typename R>
__device__ __forceinline__ R diag_of(const R (&c)[Geo<N>::RP], int j, int h) {
  constexpr int H = Geo<N>::H, RP = Geo<N>::RP;
  R d = R(0);
#pragma unroll
  for (int r = 0; r < RP; ++r) d = (h * RP + r == j) ? c[r] : d;
  if constexpr (H == 2) {
    const R other = swap_pair(d);
    d = (h == j / RP) ? d : other;
  }
  return d;
}

}  // namespace dev
}  // namespace adrt
