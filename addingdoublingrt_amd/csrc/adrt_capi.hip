// adrt_capi.hip -- C-ABI of the batched adding-doubling solver (host side).
//
// Validation mirrors adrt::ADConfig::validate (reference
// src/adding_doubling.hpp:123-174); the quadrature, Legendre and Planck
// helpers restate src/quadrature.cpp:11-76 and src/planck.cpp:35-177 in host
// C++ of this library (they are also the host half of the product: the C++
// drop-in adrt::solve uses them for thermal inputs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/adrt_hip.h"
#include "adrt_fwd.hpp"
#include "adrt_adj16.hpp"
#include "adrt_mix16.hpp"
#include "adrt_adjm16.hpp"
#include "adrt_pair16.hpp"
#include "adrt_wave32.hpp"
#include "adrt_mix32.hpp"
#include "adrt_adj32.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

#define ADRT_HIP_TRY(expr)                    \
  do {                                        \
    if ((expr) != hipSuccess) return ADRT_ERR_HIP; \
  } while (0)

// Tokens of forward solves, unique across every workspace of the process, so
// that a token kept from one workspace can never match a solve on another
// (for instance on a workspace rebuilt at the same address).
std::atomic<unsigned long long> g_forward_tokens{0};

int padded_order(int n) {
  int p = 2;
  while (p < n) p *= 2;
  return p;
}

// Gauss-Legendre nodes/weights on [0,1], ascending (src/quadrature.cpp:11-50).
void gauss_legendre01(int n, double* nodes, double* weights) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double pk = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      pp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = -p1 / pp;
      z += dz;
      if (std::fabs(dz) < 1e-15) break;
    }
    const double w = 2.0 / ((1.0 - z * z) * pp * pp);
    nodes[i] = (1.0 - z) / 2.0;
    nodes[n - 1 - i] = (1.0 + z) / 2.0;
    weights[i] = w / 2.0;
    weights[n - 1 - i] = w / 2.0;
  }
}

// P_l(x) by the three-term recurrence (src/quadrature.cpp:53-76).
void legendre_column(int L, double x, double* out) {
  if (L > 0) out[0] = 1.0;
  if (L > 1) out[1] = x;
  for (int l = 2; l < L; ++l) out[l] = ((2 * l - 1) * x * out[l - 1] - (l - 1) * out[l - 2]) / l;
}

// ---- band Planck function (src/planck.cpp:35-131) --------------------------
constexpr double kC2 = 1.438786, kSigma = 5.67032E-8, kVcut = 1.5;
constexpr double kA1 = 1.0 / 3.0, kA2 = -1.0 / 8.0, kA3 = 1.0 / 60.0, kA4 = -1.0 / 5040.0,
                 kA5 = 1.0 / 272160.0, kA6 = -1.0 / 13305600.0;

__host__ __device__ inline double plkernel(double v) { return v * v * v / (exp(v) - 1.0); }

__host__ __device__ int planck_band(double lo, double hi, double T, double* out) {
  if (T < 0.0 || hi < lo || lo < 0.0) return ADRT_ERR_PLANCK_ARGS;
  if (T < 1.e-4) { *out = 0.0; return ADRT_OK; }
  if (hi == lo) {
    const double arg = exp(-kC2 * hi / T);
    *out = 1.1911e-8 * hi * hi * hi * arg / (1.0 - arg);
    return ADRT_OK;
  }
  const double vmax = 709.782712893384;  // log(DBL_MAX)
  const double sigdpi = kSigma / kPi;
  const double conc = 15.0 / pow(kPi, 4.0);
  const double v[2] = {kC2 * lo / T, kC2 * hi / T};
  if (v[0] > DBL_EPSILON && v[1] < vmax && (hi - lo) / hi < 1.e-2) {
    // narrow band: Simpson's rule refined until 1e-6 relative change
    const double hh = v[1] - v[0];
    double oldval = 0.0;
    const double val0 = plkernel(v[0]) + plkernel(v[1]);
    for (int n = 1; n <= 10; ++n) {
      const double del = hh / (2.0 * n);
      double val = val0;
      for (int k = 1; k <= 2 * n - 1; ++k) val += 2.0 * (1 + k % 2) * plkernel(v[0] + k * del);
      val *= del * kA1;
      if (fabs((val - oldval) / val) <= 1.e-6) {
        *out = sigdpi * pow(T, 4.0) * conc * val;
        return ADRT_OK;
      }
      oldval = val;
    }
    *out = sigdpi * pow(T, 4.0) * conc * oldval;
    return ADRT_OK;
  }
  // series in v (small v) or in exp(-m v) (large v)
  double p[2] = {0.0, 0.0}, d[2] = {0.0, 0.0};
  int smallv = 0;
  const double vcp[7] = {10.25, 5.7, 3.9, 2.9, 2.3, 1.9, 0.0};
  for (int i = 0; i < 2; ++i) {
    if (v[i] < kVcut) {
      ++smallv;
      const double vsq = v[i] * v[i];
      p[i] = conc * vsq * v[i] *
             (kA1 + v[i] * (kA2 + v[i] * (kA3 + vsq * (kA4 + vsq * (kA5 + vsq * kA6)))));
    } else {
      int mmax = 1;
      while (v[i] < vcp[mmax - 1]) ++mmax;
      const double ex = exp(-v[i]);
      double exm = 1.0;
      d[i] = 0.0;
      for (int m = 1; m <= mmax; ++m) {
        const double mv = m * v[i];
        exm *= ex;
        d[i] += exm * (6.0 + mv * (6.0 + mv * (3.0 + mv))) / (m * m * m * m);
      }
      d[i] *= conc;
    }
  }
  double ans;
  if (smallv == 2) ans = p[1] - p[0];
  else if (smallv == 1) ans = 1.0 - p[0] - d[1];
  else ans = d[0] - d[1];
  ans *= sigdpi * pow(T, 4.0);
  *out = (ans == 0.0) ? 0.0 : ans;
  return ADRT_OK;
}

// dB/dT (src/planck.cpp:135-177)
__host__ __device__ int planck_band_deriv(double lo, double hi, double T, double* out) {
  if (T < 0.0 || hi < lo || lo < 0.0) return ADRT_ERR_PLANCK_ARGS;
  if (T < 1.e-4) { *out = 0.0; return ADRT_OK; }
  auto fk = [](double v) -> double {
    if (v < 1.e-6) return v * v;
    const double em = exp(-v);
    return v * v * v * em / (1.0 - em);
  };
  if (hi == lo) {
    const double x = kC2 * hi / T;
    const double em = exp(-x);
    const double b = 1.1911e-8 * hi * hi * hi * em / (1.0 - em);
    *out = b * (x / T) / (1.0 - em);
    return ADRT_OK;
  }
  const double sigdpi = kSigma / kPi;
  const double conc = 15.0 / pow(kPi, 4.0);
  const double v0 = kC2 * lo / T, v1 = kC2 * hi / T;
  double B;
  planck_band(lo, hi, T, &B);
  const double boundary = sigdpi * conc * T * T * T * (v1 * fk(v1) - v0 * fk(v0));
  *out = 4.0 * B / T - boundary;
  return ADRT_OK;
}

int check_config(const adrt_batch_config* c) {
  if (!c) return ADRT_ERR_NULL;
  if (c->num_wavenumbers < 0) return ADRT_ERR_NUM_WAVENUMBERS;
  if (c->num_layers <= 0) return ADRT_ERR_NUM_LAYERS;
  if (c->num_quadrature < 2) return ADRT_ERR_NUM_QUADRATURE;
  if (c->num_quadrature > ADRT_MAX_QUADRATURE) return ADRT_ERR_QUADRATURE_MAX;
  if (c->num_moments < 1) return ADRT_ERR_MOMENTS;
  if (!(c->surface_albedo >= 0.0 && c->surface_albedo <= 1.0)) return ADRT_ERR_SURFACE_ALBEDO;
  if (c->solar_flux > 0.0 && (c->solar_mu <= 0.0 || c->solar_mu > 1.0)) return ADRT_ERR_SOLAR_MU;
  return ADRT_OK;
}

// Planck values (or dB/dT) at every level of every column:
// out[w][l] = B(wn_lo[w], wn_hi[w], T[w][l] or T[l]).
__global__ void planck_levels_kernel(const double* lo, const double* hi, const double* T,
                                     int t_shared, int W, int nlev, int deriv, double* out,
                                     int* flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)W * nlev) return;
  const int w = (int)(i / nlev), l = (int)(i % nlev);
  const double t = t_shared ? T[l] : T[i];
  double b = 0.0;
  const int rc = deriv ? planck_band_deriv(lo[w], hi[w], t, &b) : planck_band(lo[w], hi[w], t, &b);
  if (rc) atomicOr(flag, 1);
  out[i] = b;
}

// Single-wavenumber Planck function in W m^-2 cm, the form and constants of the
// retrieval code the raw entry serves (src_cuda/cuda_planck_single.cuh:18-34):
// c1 = 2 h c^2 (CGS, scaled to W m^-2 cm), c2 = h c / k.
__device__ __forceinline__ double planck_single(double T, double nu) {
  constexpr double c1 = 1.19105e-08, c2 = 1.43879;
  if (T < 1e-5) return 0.0;
  return c1 * nu * nu * nu / expm1(c2 * nu / T);
}

// raw coefficients -> TOA-first tau, omega, level Planck and boundary emission
// for the wavenumbers w0 .. w0+W-1 of a batch of nwav (src_cuda/cuda_solver.cu:1131-1182)
__global__ void raw_optics_kernel(int W, int w0, int nwav, int L, const double* abs_c,
                                  const double* sca_c, const double* alt, const double* T,
                                  const double* wn, const double* cloud, double Ts, double Tt,
                                  double* tau_o, double* ssa_o, double* B_o, double* Bs_o,
                                  double* Bt_o) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const long gw = (long)w0 + w;
  const double nu = wn[gw];
  for (int l = 0; l < L; ++l) {
    const int b = L - 1 - l;  // BOA-first layer index
    const double dz = alt[b + 1] - alt[b];
    const double ab = abs_c[(long)b * nwav + gw], at = abs_c[(long)(b + 1) * nwav + gw];
    const double sb = sca_c[(long)b * nwav + gw], st = sca_c[(long)(b + 1) * nwav + gw];
    double tau = dz * ((at + st) + (ab + sb)) * 0.5;
    const double sdepth = dz * (st + sb) * 0.5;
    if (cloud) tau += cloud[(long)b * nwav + gw];
    if (tau < 0.0) tau = 0.0;
    tau_o[(long)w * L + l] = tau;
    ssa_o[(long)w * L + l] = tau > 0.0 ? sdepth / tau : 0.0;
  }
  for (int lev = 0; lev <= L; ++lev) B_o[(long)w * (L + 1) + lev] = planck_single(T[L - lev], nu);
  Bs_o[w] = planck_single(Ts >= 0.0 ? Ts : T[0], nu);
  Bt_o[w] = planck_single(Tt >= 0.0 ? Tt : T[L], nu);
}

// dB/dT of planck_single: B (x / T) e^x / (e^x - 1), x = c2 nu / T
__device__ __forceinline__ double planck_single_dT(double T, double nu) {
  constexpr double c1 = 1.19105e-08, c2 = 1.43879;
  if (T < 1e-5) return 0.0;
  const double x = c2 * nu / T;
  const double em = expm1(x);
  return c1 * nu * nu * nu / em * (x / T) * (1.0 + 1.0 / em);
}

// reverse of raw_optics_kernel for tau and omega: cotangents of the TOA-first
// [W][L] optics to the level-major absorption / scattering coefficients and
// the cloud depths (thread per wavenumber; each level collects its two layers
// in order, no atomics).  A clamped layer (tau <= 0) passes nothing back.
__global__ void raw_optics_adjoint_kernel(int W, int L, const double* abs_c, const double* sca_c,
                                          const double* alt, const double* cloud,
                                          const double* g_tau, const double* g_ssa,
                                          double* g_abs, double* g_sca, double* g_cloud) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  double ga_lo = 0.0, gs_lo = 0.0;  // running cotangent of level b (from layer b - 1)
  for (int b = 0; b < L; ++b) {
    const int l = L - 1 - b;
    const double dz = alt[b + 1] - alt[b];
    const double ab = abs_c[(long)b * W + w], at = abs_c[(long)(b + 1) * W + w];
    const double sb = sca_c[(long)b * W + w], st = sca_c[(long)(b + 1) * W + w];
    double tau = dz * ((at + st) + (ab + sb)) * 0.5;
    const double sdepth = dz * (st + sb) * 0.5;
    if (cloud) tau += cloud[(long)b * W + w];
    double gt = 0.0, gsd = 0.0;
    if (tau > 0.0) {
      const double gtau = g_tau ? g_tau[(long)w * L + l] : 0.0;
      const double gom = g_ssa ? g_ssa[(long)w * L + l] : 0.0;
      gt = gtau - gom * sdepth / (tau * tau);  // omega = sdepth / tau
      gsd = gom / tau;
    }
    if (g_cloud) g_cloud[(long)b * W + w] = gt;
    const double ge = gt * dz * 0.5, gs = (gt + gsd) * dz * 0.5;  // per level of the pair
    if (g_abs) g_abs[(long)b * W + w] = ga_lo + ge;
    if (g_sca) g_sca[(long)b * W + w] = gs_lo + gs;
    ga_lo = ge;
    gs_lo = gs;
  }
  if (g_abs) g_abs[(long)L * W + w] = ga_lo;
  if (g_sca) g_sca[(long)L * W + w] = gs_lo;
}

// reverse of the Planck values: block k < L+1 sums over the wavenumbers the
// cotangent of level k (BOA first; TOA-first level L - k) times dB/dT, plus
// the surface / top emission where they follow the bottom / top level; blocks
// L+1 and L+2 give d/d surface_temperature and d/d top_temperature when those
// are set.  Fixed-order sums: deterministic.
__global__ void raw_temperature_adjoint_kernel(int W, int L, const double* T, const double* wn,
                                               double Ts, double Tt, const double* g_B,
                                               const double* g_Bs, const double* g_Bt,
                                               double* g_T, double* g_Ts, double* g_Tt) {
  __shared__ double part[256];
  const int k = blockIdx.x;
  double acc = 0.0;
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    const double nu = wn[w];
    if (k <= L) {
      const double d = planck_single_dT(T[k], nu);
      double g = g_B ? g_B[(long)w * (L + 1) + (L - k)] : 0.0;
      if (k == 0 && Ts < 0.0 && g_Bs) g += g_Bs[w];
      if (k == L && Tt < 0.0 && g_Bt) g += g_Bt[w];
      acc += g * d;
    } else if (k == L + 1) {
      if (g_Bs) acc += g_Bs[w] * planck_single_dT(Ts, nu);
    } else {
      if (g_Bt) acc += g_Bt[w] * planck_single_dT(Tt, nu);
    }
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (k <= L) {
      if (g_T) g_T[k] = part[0];
    } else if (k == L + 1) {
      if (g_Ts) g_Ts[0] = Ts >= 0.0 ? part[0] : 0.0;
    } else {
      if (g_Tt) g_Tt[0] = Tt >= 0.0 ? part[0] : 0.0;
    }
  }
}

__global__ void validate_kernel(const double* tau, const double* ssa, long n, int* flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (tau[i] < 0.0) atomicOr(flag, 1);
  const double w = ssa[i];
  if (w < 0.0 || w > 1.0) atomicOr(flag, 2);
}

}  // namespace

struct adrt_workspace {
  int L = 0, n_real = 0, Np = 0, M = 0, Lmax = 0;
  int chunk = 0;
  size_t bytes = 0;
  int device = 0;
  char* base = nullptr;
  double *mu = nullptr, *wt = nullptr, *Pl = nullptr, *Pl0 = nullptr;
  double *tau_s = nullptr, *omega_s = nullptr, *tau_cum = nullptr, *fdm = nullptr;
  double *lay = nullptr, *rb = nullptr;
  int *mode = nullptr, *lflag = nullptr, *flag = nullptr, *redo = nullptr, *ucol = nullptr;
  int *ulist = nullptr, *ucount = nullptr;  // the flagged columns (FwdArgs::ulist)
  // reverse pass (ADRT_WORKSPACE_ADJOINT)
  int flags = 0;
  int *pord = nullptr, *pflag = nullptr;  // pair mode (orders <= 8 padded to 16)
  double *Iv = nullptr, *Av = nullptr, *g_taus = nullptr, *g_omgs = nullptr, *g_Bt = nullptr,
         *g_Bb = nullptr, *g_Ftop = nullptr, *g_f = nullptr, *g_chi = nullptr, *scot = nullptr, *slab = nullptr,
         *mom_part = nullptr, *phase = nullptr, *raw = nullptr;
  int* status = nullptr;
  int* queue = nullptr;  // task counter of the persistent reverse grids (after status)
  // source Jacobians (ADRT_WORKSPACE_JACOBIAN): the thermal basis of the
  // chunk's columns and (L + 3) pseudo-columns per column
  double *basis = nullptr, *jlay = nullptr, *jrb = nullptr, *jtau_s = nullptr,
         *jtau_cum = nullptr, *jB = nullptr, *jsurf = nullptr, *jtop = nullptr, *jssa = nullptr,
         *jfdir = nullptr;
  int *jlflag = nullptr, *jucol = nullptr;
  int nwaves = 0, nn_cap = 0, cus = 0;  // slabs (waves of the larger reverse grid); CUs
  unsigned long long generation = 0;  // token of the last forward solve (unique in the process)
  bool kept_valid = false;            // the last forward's state (one chunk) is still held
  bool kept_mixed = false;            // ... and was computed by the mixed-precision form
  double xfac = 0.0;
  double tables_mu0 = -1.0;
  bool tables_ready = false;
  std::vector<double> h_mu, h_wt;
  // opt-in kernel timing: event pairs recorded on the caller's stream around
  // every launch (kinds as adrt_profile_read)
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  // the composite-form adding of the flagged columns runs on this stream beside
  // the base sweep on the caller's stream (fork / join by the two events);
  // null for a workspace that has none (then both run on the caller's stream)
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::vector<int> ev_kind;  // one per recorded pair
  size_t ev_used = 0;        // pairs in use
};

namespace {

size_t per_wavenumber_bytes(int L, int Np, int M, bool adjoint) {
  const size_t d = sizeof(double);
  size_t b = 0;
  b += 4 * (size_t)L * d + (size_t)(L + 1) * d;      // tau_s, omega_s, fdm, spare, tau_cum
  b += 3 * (size_t)L * sizeof(int) + 2 * sizeof(int); // mode, lflag, redo; ucol, ulist
  b += (size_t)L * (2 * Np * Np + 2 * Np) * d;        // layer store
  b += (size_t)L * (2 * Np * Np + 4 * Np) * d;        // adding records
  b += (size_t)(3 * L + 3) * d;                       // raw-coefficient optics
  if (adjoint) {
    b += 2 * (size_t)(L + 1) * 2 * Np * d;            // forward and adjoint interface vectors
    b += 6 * (size_t)L * d;                           // per-layer gradient terms
    b += 2 * (size_t)L * sizeof(int);                 // pair order and flags (orders <= 8)
    b += (size_t)L * M * d;                           // per-layer moment gradients
    if (Np == adrt::dev::w16::N)
      b += (size_t)L * adrt::dev::w16::SCOT * d;      // doubling-start cotangents
    else if (Np == adrt::dev::w32::N)
      b += (size_t)L * adrt::dev::w32::SCOT32 * d;
  }
  return b;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int upload_tables(adrt_workspace* ws, double mu0, hipStream_t stream) {
  if (ws->tables_ready && ws->tables_mu0 == mu0) return ADRT_OK;
  const int Np = ws->Np, n = ws->n_real, Lmax = ws->Lmax;
  std::vector<double> h((size_t)(2 * Np + Lmax * Np + Lmax), 0.0);
  double* mu = h.data();
  double* wt = mu + Np;
  double* Pl = wt + Np;
  double* Pl0 = Pl + (size_t)Lmax * Np;
  for (int i = 0; i < Np; ++i) {
    mu[i] = (i < n) ? ws->h_mu[i] : 1.0;  // pad nodes: mu = 1, w = 0, P_l = 0
    wt[i] = (i < n) ? ws->h_wt[i] : 0.0;
  }
  std::vector<double> col((size_t)Lmax);
  for (int i = 0; i < n; ++i) {
    legendre_column(Lmax, mu[i], col.data());
    for (int l = 0; l < Lmax; ++l) Pl[(size_t)l * Np + i] = col[l];
  }
  legendre_column(Lmax, mu0, Pl0);  // P_l(mu0), src/phase_matrix.hpp:145-151
  ADRT_HIP_TRY(hipMemcpyAsync(ws->mu, h.data(), h.size() * sizeof(double),
                              hipMemcpyHostToDevice, stream));
  ADRT_HIP_TRY(hipStreamSynchronize(stream));
  ws->tables_ready = true;
  ws->tables_mu0 = mu0;
  return ADRT_OK;
}

// Records the start event of a timed launch; returns the pair index or -1.
int prof_begin(adrt_workspace* ws, int kind, hipStream_t stream) {
  if (!ws->profiling) return -1;
  if (ws->ev_used * 2 + 2 > ws->ev_pool.size()) {
    for (int i = 0; i < 2; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      ws->ev_pool.push_back(e);
    }
    ws->ev_kind.push_back(kind);
  }
  const size_t k = ws->ev_used++;
  ws->ev_kind[k] = kind;
  (void)hipEventRecord(ws->ev_pool[2 * k], stream);  // profiling only
  return (int)k;
}

void prof_end(adrt_workspace* ws, int k, hipStream_t stream) {
  if (k >= 0) (void)hipEventRecord(ws->ev_pool[2 * k + 1], stream);
}

// prep, phase records and the layer kernel (layer records) of one chunk
template <int N>
int layers_chunk(const adrt::dev::FwdArgs& a, adrt_workspace* ws, hipStream_t stream,
                 bool mixed) {
  using namespace adrt::dev;
  constexpr int GPB = Geo<N>::GPB;  // lane groups per 256-thread block
  int e = prof_begin(ws, 0, stream);
  if (a.ulist) w16::fill_int_kernel<<<1, 1, 0, stream>>>(a.ucount, 0);  // a kernel: capturable
  prep_kernel<<<(a.W + 255) / 256, 256, 0, stream>>>(a);
  prof_end(ws, e, stream);
  const long tasks = (long)a.W * a.L;
  e = prof_begin(ws, 1, stream);
  if constexpr (N == 16) {  // one wave per problem on the matrix cores
    if (a.phase) w16::phase16_kernel<<<(unsigned)a.L, 64, 0, stream>>>(a, ws->phase);
    if (mixed) {
      w16::layer16m_kernel<<<(unsigned)((tasks + 3) / 4), 256, 0, stream>>>(a);
      // layers whose doubling start is outside FP32's reach, again in FP64
      w16::layer16_kernel<true><<<(unsigned)((tasks + 3) / 4), 256, 0, stream>>>(a);
    } else if (a.pord) {
      // orders up to 8 padded to 16 (reverse-pass workspaces): two layers per wave
      const long pairs = (long)a.W * (a.L / 2);
      w16::pair_order_kernel<<<(unsigned)((a.W + 3) / 4), 256, 0, stream>>>(a);
      if (pairs > 0) w16::layer16p_kernel<<<(unsigned)((pairs + 3) / 4), 256, 0, stream>>>(a);
      w16::layer16_kernel<false, true><<<(unsigned)((tasks + 3) / 4), 256, 0, stream>>>(a);
    } else {
      w16::layer16_kernel<false><<<(unsigned)((tasks + 3) / 4), 256, 0, stream>>>(a);
    }
  } else {
    if constexpr (N == 32) {  // one wave per problem on the matrix cores
      const unsigned g = (unsigned)((tasks + w32::WAVES - 1) / w32::WAVES);
      if (mixed) {
        w32::layer32m_kernel<<<g, 64 * w32::WAVES, 0, stream>>>(a);
        // the routed and the non-scattering layers, in FP64
        w32::layer32_kernel<true><<<g, 64 * w32::WAVES, 0, stream>>>(a);
      } else {
        w32::layer32_kernel<false><<<g, 64 * w32::WAVES, 0, stream>>>(a);
      }
    } else {
      if (mixed) return ADRT_ERR_PRECISION;
      layer_kernel<N><<<(unsigned)((tasks + GPB - 1) / GPB), 256, 0, stream>>>(a);
    }
  }
  prof_end(ws, e, stream);
  return hipGetLastError() == hipSuccess ? ADRT_OK : ADRT_ERR_HIP;
}

// The adding of a chunk runs as two kernels over disjoint columns: the base
// sweep, and the reference's composite form for the columns the prep kernel
// flagged.  Each wave sweeps one column's layers in sequence, so a launch lasts
// about one column's latency however few columns it has; the composite
// kernel therefore runs beside the base sweep, on the workspace's side stream,
// forked from and joined back into the caller's stream by events.
template <class Base, class Side>
int fork_join(adrt_workspace* ws, hipStream_t stream, Base base, Side side) {
  if (!ws->side) {
    base(stream);
    side(stream);
    return ADRT_OK;
  }
  if (hipEventRecord(ws->fork, stream) != hipSuccess ||
      hipStreamWaitEvent(ws->side, ws->fork, 0) != hipSuccess)
    return ADRT_ERR_HIP;
  side(ws->side);
  base(stream);
  if (hipEventRecord(ws->join, ws->side) != hipSuccess ||
      hipStreamWaitEvent(stream, ws->join, 0) != hipSuccess)
    return ADRT_ERR_HIP;
  return ADRT_OK;
}

// adding sweeps, interface solves and outputs of one chunk
template <int N>
int adding_chunk(const adrt::dev::FwdArgs& a, adrt_workspace* ws, hipStream_t stream,
                 bool mixed) {
  using namespace adrt::dev;
  constexpr int GPB = Geo<N>::GPB;
  const int e = prof_begin(ws, 2, stream);
  int rc;
  if constexpr (N == 16) {
    // the base sweep; columns flagged by the prep kernel take the composite form
    const unsigned g = (unsigned)((a.W + 3) / 4);
    if (mixed)
      rc = fork_join(
          ws, stream,
          [&](hipStream_t s) { w16::adding16_kernel<false, true><<<g, 256, 0, s>>>(a); },
          [&](hipStream_t s) { w16::adding16_composite_kernel<false, true><<<g, 256, 0, s>>>(a); });
    else
      rc = fork_join(
          ws, stream,
          [&](hipStream_t s) { w16::adding16_kernel<false, false><<<g, 256, 0, s>>>(a); },
          [&](hipStream_t s) { w16::adding16_composite_kernel<false, false><<<g, 256, 0, s>>>(a); });
  } else if constexpr (N == 32) {
    const unsigned g = (unsigned)((a.W + w32::WAVES - 1) / w32::WAVES);
    rc = fork_join(
        ws, stream,
        [&](hipStream_t s) { w32::adding32_kernel<false><<<g, 64 * w32::WAVES, 0, s>>>(a); },
        [&](hipStream_t s) { w32::adding32_composite_kernel<false><<<g, 64 * w32::WAVES, 0, s>>>(a); });
  } else {
    const unsigned g = (unsigned)((a.W + GPB - 1) / GPB);
    rc = fork_join(
        ws, stream, [&](hipStream_t s) { adding_kernel<N><<<g, 256, 0, s>>>(a); },
        [&](hipStream_t s) { adding_composite_kernel<N><<<g, 256, 0, s>>>(a); });
  }
  prof_end(ws, e, stream);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? ADRT_OK : ADRT_ERR_HIP;
}

template <int N>
int launch_chunk(const adrt::dev::FwdArgs& a, adrt_workspace* ws, hipStream_t stream,
                 bool mixed) {
  const int rc = layers_chunk<N>(a, ws, stream, mixed);
  return rc ? rc : adding_chunk<N>(a, ws, stream, mixed);
}

// Source Jacobians: every output is affine in the thermal sources and R, T do
// not depend on them, so d(output)/d(source u) is the solve with source u set
// to one and every other source (and the beam) zero.  Pseudo-column
// p = c (L + 3) + u of real column c takes the real column's R, T and its
// sources rebuilt from the thermal basis for unit source u: u <= L the Planck
// value of level u (caller order), L + 1 the surface emission, L + 2 the top.
struct JacArgs {
  int C, L, Np, rev;
  const double* lay;    // [C][L][REC] real records
  const double* basis;  // [C][L][2 Np]
  const int* lflag;     // [C][L]
  const double* tau_s;  // [C][L]
  const double* tau_cum;  // [C][L+1]
  const double* ssa;    // [caller columns][L] (the chunk's columns from w0), caller order
  long w0;
  double* jlay;         // [P][L][REC]
  int* jlflag;          // [P][L]
  const int* ucol;      // [C]
  int* jucol;           // [P]
  double *jtau_s, *jtau_cum, *jB, *jsurf, *jtop, *jssa;
};

__global__ void jac_records_kernel(JacArgs j) {
  const long task = blockIdx.x;  // pseudo-column * L + layer
  const int U = j.L + 3;
  const long p = task / j.L;
  const int l = (int)(task % j.L);
  const long c = p / U;
  const int u = (int)(p % U);
  const int REC = 2 * j.Np * j.Np + 2 * j.Np;
  const double* src = j.lay + (c * j.L + l) * REC;
  double* dst = j.jlay + task * REC;
  for (int k = threadIdx.x; k < 2 * j.Np * j.Np; k += blockDim.x) dst[k] = src[k];
  const int e = threadIdx.x;
  if (e < j.Np) {
    const int lt = j.rev ? j.L - l : l, lb = j.rev ? j.L - l - 1 : l + 1;  // caller levels
    const double Bt = (u == lt) ? 1.0 : 0.0, Bb = (u == lb) ? 1.0 : 0.0;
    const double tau = j.tau_s[c * j.L + l];
    const double B_bar = (Bb + Bt) / 2.0;
    const double B_d = (tau > 0.0) ? (Bb - Bt) / tau : 0.0;
    const double* bs = j.basis + (c * j.L + l) * 2 * j.Np;
    const double y = bs[e], z = bs[j.Np + e];
    dst[2 * j.Np * j.Np + e] = y * B_bar + z * B_d;
    dst[2 * j.Np * j.Np + j.Np + e] = y * B_bar - z * B_d;
  }
}

__global__ void jac_columns_kernel(JacArgs j, long P) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int U = j.L + 3, L = j.L;
  const long c = p / U;
  const int u = (int)(p % U);
  for (int l = 0; l < L; ++l) {
    j.jlflag[p * L + l] = j.lflag[c * L + l];
    j.jtau_s[p * L + l] = j.tau_s[c * L + l];
    if (j.jssa) j.jssa[p * L + l] = j.ssa[(j.w0 + c) * L + l];
  }
  for (int l = 0; l <= L; ++l) {
    j.jtau_cum[p * (L + 1) + l] = j.tau_cum[c * (L + 1) + l];
    j.jB[p * (L + 1) + l] = (u == l) ? 1.0 : 0.0;
  }
  j.jucol[p] = j.ucol[c];
  j.jsurf[p] = (u == L + 1) ? 1.0 : 0.0;
  j.jtop[p] = (u == L + 2) ? 1.0 : 0.0;
}

}  // namespace

extern "C" {

int adrt_workspace_create(const adrt_batch_config* cfg, size_t max_bytes, adrt_workspace** out) {
  return adrt_workspace_create_ex(cfg, max_bytes, 0, out);
}

int adrt_workspace_create_ex(const adrt_batch_config* cfg, size_t max_bytes, int flags,
                             adrt_workspace** out) {
  if (!out) return ADRT_ERR_NULL;
  *out = nullptr;
  int rc = check_config(cfg);
  if (rc) return rc;
  const bool adjoint = (flags & ADRT_WORKSPACE_ADJOINT) != 0;
  const bool jac = (flags & ADRT_WORKSPACE_JACOBIAN) != 0;
  if (adjoint && jac) return ADRT_ERR_WORKSPACE;
  adrt_workspace* ws = new (std::nothrow) adrt_workspace();
  if (!ws) return ADRT_ERR_WORKSPACE;
  ws->flags = flags;
  ws->L = cfg->num_layers;
  ws->n_real = cfg->num_quadrature;
  // reverse-pass workspaces: orders up to 16 padded to 16, 17..32 to 32
  ws->Np = adjoint ? (cfg->num_quadrature <= 16 ? 16 : 32) : padded_order(cfg->num_quadrature);
  ws->M = cfg->num_moments;
  ws->Lmax = std::max(cfg->num_moments, 2 * cfg->num_quadrature);
  ws->h_mu.assign(ws->n_real, 0.0);
  ws->h_wt.assign(ws->n_real, 0.0);
  gauss_legendre01(ws->n_real, ws->h_mu.data(), ws->h_wt.data());
  double s = 0.0;
  for (int i = 0; i < ws->n_real; ++i) s += ws->h_mu[i] * ws->h_wt[i];
  ws->xfac = 0.5 / s;  // src/solver.cpp:378-383
  if (hipGetDevice(&ws->device) != hipSuccess) {
    delete ws;
    return ADRT_ERR_HIP;
  }
  size_t slab_bytes = 0;
  if (adjoint) {
    // persistent reverse-layer grids: REV_WAVES_PER_SIMD (FP64) or
    // MREV_WAVES_PER_SIMD (FP32) waves on each of 4 SIMDs per CU; one slab
    // per wave of the larger grid
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ws->device) !=
            hipSuccess || cus <= 0)
      cus = 256;
    ws->cus = cus;
    if (ws->Np == adrt::dev::w32::N) {
      // one wave per SIMD; 48 steps hold layers up to tau 2^32
      ws->nwaves = cus * 4 * adrt::dev::w32::REV32_WAVES_PER_SIMD;
      ws->nn_cap = 48;
      slab_bytes = (size_t)ws->nwaves * adrt::dev::w32::slab_stride32(ws->nn_cap) * sizeof(double);
    } else {
      ws->nwaves = cus * 4 * std::max(adrt::dev::w16::REV_WAVES_PER_SIMD,
                                      adrt::dev::w16::MREV_WAVES_PER_SIMD);
      ws->nn_cap = 64;
      slab_bytes = (size_t)ws->nwaves * adrt::dev::w16::slab_stride(ws->nn_cap) * sizeof(double);
    }
  }
  if (max_bytes == 0) max_bytes = (size_t)4 << 30;
  size_t per_w = per_wavenumber_bytes(ws->L, ws->Np, ws->M, adjoint);
  if (jac) {  // thermal basis + L + 3 pseudo-columns per column
    const size_t Lj = ws->L, Nj = ws->Np;
    const size_t pseudo = Lj * (2 * Nj * Nj + 2 * Nj) * 8 + Lj * (2 * Nj * Nj + Nj) * 8 +
                          Lj * 4 + 2 * Lj * 8 + 3 * (Lj + 1) * 8 + 2 * 8;
    per_w += Lj * 2 * Nj * 8 + (Lj + 3) * pseudo;
  }
  const size_t phase_bytes =
      ws->Np == adrt::dev::w16::N ? (size_t)ws->L * adrt::dev::w16::PH16 * 8 : 0;
  const size_t fixed =
      align256((size_t)(2 * ws->Np + ws->Lmax * ws->Np + ws->Lmax) * 8) + 768 + slab_bytes +
      align256(phase_bytes) +
      (adjoint ? align256((size_t)ws->L * adrt::dev::w16::MOM_SLICES * ws->M * 8) : 0);
  long chunk = max_bytes > fixed ? (long)((max_bytes - fixed) / (per_w + 4096)) : 1;
  const int W = std::max(cfg->num_wavenumbers, 1);
  chunk = std::max(1L, std::min(chunk, (long)W));
  ws->chunk = (int)chunk;
  const size_t L = ws->L, Np = ws->Np, C = chunk, M = ws->M;
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += align256(b);
    return o;
  };
  const size_t o_tab = take((size_t)(2 * Np + ws->Lmax * Np + ws->Lmax) * 8);
  const size_t o_tau_s = take(C * L * 8), o_om = take(C * L * 8), o_fdm = take(C * L * 8);
  const size_t o_cum = take(C * (L + 1) * 8);
  const size_t o_mode = take(C * L * 4), o_lflag = take(C * L * 4), o_flag = take(4);
  const size_t o_redo = take(C * L * 4), o_ucol = take(C * 4), o_ulist = take(C * 4);
  const size_t o_ucount = take(4);
  const size_t o_lay = take(C * L * (2 * Np * Np + 2 * Np) * 8);
  const size_t o_rb = take(C * L * (2 * Np * Np + 4 * Np) * 8);
  const size_t o_phase = take(phase_bytes);
  const size_t o_raw = take(C * (3 * L + 3) * 8);
  size_t o_iv = 0, o_av = 0, o_g = 0, o_chi = 0, o_slab = 0, o_status = 0, o_part = 0,
         o_scot = 0, o_pord = 0, o_pflag = 0;
  if (adjoint) {
    o_iv = take(C * (L + 1) * 2 * Np * 8);
    o_av = take(C * (L + 1) * 2 * Np * 8);
    o_g = take(6 * C * L * 8);
    o_pord = take(C * L * 4);
    o_pflag = take(C * L * 4);
    o_part = take(L * adrt::dev::w16::MOM_SLICES * M * 8);
    o_chi = take(C * L * M * 8);
    o_scot = take(Np == adrt::dev::w16::N   ? C * L * adrt::dev::w16::SCOT * 8
                  : Np == adrt::dev::w32::N ? C * L * adrt::dev::w32::SCOT32 * 8
                                            : 0);
    o_slab = take(slab_bytes);
    o_status = take(8);  // status, queue
  }
  size_t o_basis = 0, o_jlay = 0, o_jrb = 0, o_jlflag = 0, o_jtau_s = 0, o_jtau_cum = 0,
         o_jB = 0, o_jsurf = 0, o_jtop = 0, o_jssa = 0, o_jfdir = 0, o_jucol = 0;
  if (jac) {
    const size_t P = C * (L + 3);
    o_basis = take(C * L * 2 * Np * 8);
    o_jlay = take(P * L * (2 * Np * Np + 2 * Np) * 8);
    o_jrb = take(P * L * (2 * Np * Np + 4 * Np) * 8);
    o_jlflag = take(P * L * 4);
    o_jucol = take(P * 4);
    o_jtau_s = take(P * L * 8);
    o_jtau_cum = take(P * (L + 1) * 8);
    o_jB = take(P * (L + 1) * 8);
    o_jsurf = take(P * 8);
    o_jtop = take(P * 8);
    o_jssa = take(P * L * 8);
    o_jfdir = take(P * (L + 1) * 8);
  }
  if (hipMalloc((void**)&ws->base, off) != hipSuccess) {
    delete ws;
    return ADRT_ERR_HIP;
  }
  ws->bytes = off;
  ws->mu = (double*)(ws->base + o_tab);
  ws->wt = ws->mu + Np;
  ws->Pl = ws->wt + Np;
  ws->Pl0 = ws->Pl + (size_t)ws->Lmax * Np;
  ws->tau_s = (double*)(ws->base + o_tau_s);
  ws->omega_s = (double*)(ws->base + o_om);
  ws->fdm = (double*)(ws->base + o_fdm);
  ws->tau_cum = (double*)(ws->base + o_cum);
  ws->mode = (int*)(ws->base + o_mode);
  ws->lflag = (int*)(ws->base + o_lflag);
  ws->redo = (int*)(ws->base + o_redo);
  ws->ucol = (int*)(ws->base + o_ucol);
  ws->ulist = (int*)(ws->base + o_ulist);
  ws->ucount = (int*)(ws->base + o_ucount);
  ws->flag = (int*)(ws->base + o_flag);
  ws->lay = (double*)(ws->base + o_lay);
  ws->rb = (double*)(ws->base + o_rb);
  ws->phase = phase_bytes ? (double*)(ws->base + o_phase) : nullptr;
  ws->raw = (double*)(ws->base + o_raw);
  if (jac) {
    ws->basis = (double*)(ws->base + o_basis);
    ws->jlay = (double*)(ws->base + o_jlay);
    ws->jrb = (double*)(ws->base + o_jrb);
    ws->jlflag = (int*)(ws->base + o_jlflag);
    ws->jucol = (int*)(ws->base + o_jucol);
    ws->jtau_s = (double*)(ws->base + o_jtau_s);
    ws->jtau_cum = (double*)(ws->base + o_jtau_cum);
    ws->jB = (double*)(ws->base + o_jB);
    ws->jsurf = (double*)(ws->base + o_jsurf);
    ws->jtop = (double*)(ws->base + o_jtop);
    ws->jssa = (double*)(ws->base + o_jssa);
    ws->jfdir = (double*)(ws->base + o_jfdir);
  }
  if (adjoint) {
    ws->Iv = (double*)(ws->base + o_iv);
    ws->Av = (double*)(ws->base + o_av);
    if (ws->n_real <= 8 && ws->L <= adrt::dev::w16::PAIR_MAX_L) {
      ws->pord = (int*)(ws->base + o_pord);
      ws->pflag = (int*)(ws->base + o_pflag);
    }
    double* g = (double*)(ws->base + o_g);
    ws->g_taus = g;
    ws->g_omgs = g + C * L;
    ws->g_Bt = g + 2 * C * L;
    ws->g_Bb = g + 3 * C * L;
    ws->g_Ftop = g + 4 * C * L;
    ws->g_f = g + 5 * C * L;
    ws->g_chi = (double*)(ws->base + o_chi);
    ws->scot = (double*)(ws->base + o_scot);
    ws->mom_part = (double*)(ws->base + o_part);
    ws->slab = (double*)(ws->base + o_slab);
    ws->status = (int*)(ws->base + o_status);
    ws->queue = ws->status + 1;
    if (hipMemset(ws->status, 0, sizeof(int)) != hipSuccess) {
      (void)hipFree(ws->base);  // already failing
      delete ws;
      return ADRT_ERR_HIP;
    }
  }
  if (hipStreamCreateWithFlags(&ws->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ws->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ws->join, hipEventDisableTiming) != hipSuccess) {
    (void)adrt_workspace_destroy(ws);  // already failing
    return ADRT_ERR_HIP;
  }
  *out = ws;
  return ADRT_OK;
}

int adrt_workspace_destroy(adrt_workspace* ws) {
  if (!ws) return ADRT_OK;
  int rc = ADRT_OK;
  for (hipEvent_t e : ws->ev_pool) (void)hipEventDestroy(e);
  if (ws->fork) (void)hipEventDestroy(ws->fork);
  if (ws->join) (void)hipEventDestroy(ws->join);
  if (ws->side) (void)hipStreamDestroy(ws->side);
  if (ws->base && hipFree(ws->base) != hipSuccess) rc = ADRT_ERR_HIP;
  delete ws;
  return rc;
}

int adrt_workspace_info(const adrt_workspace* ws, size_t* bytes, int* chunk) {
  if (!ws) return ADRT_ERR_NULL;
  if (bytes) *bytes = ws->bytes;
  if (chunk) *chunk = ws->chunk;
  return ADRT_OK;
}

}  // extern "C"

namespace {

adrt::dev::FwdArgs make_args(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                             const adrt_device_outputs* out, adrt_workspace* ws, bool has_solar) {
  adrt::dev::FwdArgs a{};
  a.L = cfg->num_layers;
  a.M = cfg->num_moments;
  a.n_real = cfg->num_quadrature;
  a.shared_moments = in->phase_moments_shared ? 1 : 0;
  a.delta_m = cfg->use_delta_m ? 1 : 0;
  a.diffusion = cfg->use_diffusion_lower_bc ? 1 : 0;
  a.rev = cfg->index_from_bottom ? 1 : 0;
  a.has_solar = has_solar ? 1 : 0;
  a.albedo = cfg->surface_albedo;
  a.surf_em = cfg->surface_emission;
  a.top_em = cfg->top_emission;
  a.F0 = cfg->solar_flux;
  a.mu0 = cfg->solar_mu;
  a.scaling = cfg->spectrum_scaling;
  a.xfac = ws->xfac;
  a.tau = in->delta_tau;
  a.ssa = in->single_scat_albedo;
  a.mom = in->phase_moments;
  a.B = in->planck_levels;
  a.surf_w = in->surface_emission;
  a.top_w = in->top_emission;
  a.fu = out->flux_up;
  a.fd = out->flux_down;
  a.mi = out->mean_intensity;
  a.fdir = out->flux_direct;
  a.fdiv = out->flux_divergence;
  a.mu = ws->mu;
  a.wt = ws->wt;
  a.Pl = ws->Pl;
  a.Pl0 = ws->Pl0;
  a.tau_s = ws->tau_s;
  a.omega_s = ws->omega_s;
  a.tau_cum = ws->tau_cum;
  a.fdm = ws->fdm;
  a.mode = ws->mode;
  a.lay = ws->lay;
  a.lflag = ws->lflag;
  a.redo = ws->redo;
  a.ucol = ws->ucol;
  a.ulist = ws->ulist;
  a.ucount = ws->ucount;
  a.rb = ws->rb;
  a.pord = ws->pord;  // two layers per wave (orders <= 8 on reverse-pass workspaces)
  a.pflag = ws->pflag;
  a.Iv = ws->Iv;  // kept for the reverse pass (adjoint workspaces only)
  a.Av = ws->Av;
  // one phase record per layer when every wavenumber shares the moments
  a.phase = (a.shared_moments && ws->phase) ? ws->phase : nullptr;
  return a;
}

int forward_chunk(const adrt::dev::FwdArgs& a, adrt_workspace* ws, hipStream_t stream,
                  bool mixed) {
  switch (ws->Np) {
#ifndef ADRT_ONLY_N
    case 2: return launch_chunk<2>(a, ws, stream, mixed);
    case 4: return launch_chunk<4>(a, ws, stream, mixed);
    case 8: return launch_chunk<8>(a, ws, stream, mixed);
    case 16: return launch_chunk<16>(a, ws, stream, mixed);
    case 32: return launch_chunk<32>(a, ws, stream, mixed);
#else
    case ADRT_ONLY_N: return launch_chunk<ADRT_ONLY_N>(a, ws, stream, mixed);
#endif
    default: return ADRT_ERR_QUADRATURE_MAX;
  }
}

int layers_dispatch(const adrt::dev::FwdArgs& a, adrt_workspace* ws, hipStream_t stream) {
  switch (ws->Np) {
#ifndef ADRT_ONLY_N
    case 2: return layers_chunk<2>(a, ws, stream, false);
    case 4: return layers_chunk<4>(a, ws, stream, false);
    case 8: return layers_chunk<8>(a, ws, stream, false);
    case 16: return layers_chunk<16>(a, ws, stream, false);
    case 32: return layers_chunk<32>(a, ws, stream, false);
#else
    case ADRT_ONLY_N: return layers_chunk<ADRT_ONLY_N>(a, ws, stream, false);
#endif
    default: return ADRT_ERR_QUADRATURE_MAX;
  }
}

int adding_dispatch(const adrt::dev::FwdArgs& a, adrt_workspace* ws, hipStream_t stream) {
  switch (ws->Np) {
#ifndef ADRT_ONLY_N
    case 2: return adding_chunk<2>(a, ws, stream, false);
    case 4: return adding_chunk<4>(a, ws, stream, false);
    case 8: return adding_chunk<8>(a, ws, stream, false);
    case 16: return adding_chunk<16>(a, ws, stream, false);
    case 32: return adding_chunk<32>(a, ws, stream, false);
#else
    case ADRT_ONLY_N: return adding_chunk<ADRT_ONLY_N>(a, ws, stream, false);
#endif
    default: return ADRT_ERR_QUADRATURE_MAX;
  }
}

// the reverse pass of one chunk at quadrature order 32 (the forward's records
// and interface intensities are in the workspace): the transposed adding (base
// sweep, and the composite form beside it for the flagged columns), the
// doubling reverse on a persistent grid, the start reverse, the column terms
// and the shared-moment sums
int adjoint_chunk32(const adrt::dev::FwdArgs& a, adrt::dev::w16::AdjArgs b, adrt_workspace* ws,
                    hipStream_t stream, bool shared_moments) {
  using namespace adrt::dev;
  int e = prof_begin(ws, 3, stream);
  const unsigned g = (unsigned)((a.W + w32::WAVES - 1) / w32::WAVES);
  int rc = fork_join(
      ws, stream,
      [&](hipStream_t s) { w32::adding32_kernel<true><<<g, 64 * w32::WAVES, 0, s>>>(a); },
      [&](hipStream_t s) { w32::adding32_composite_kernel<true><<<g, 64 * w32::WAVES, 0, s>>>(a); });
  prof_end(ws, e, stream);
  if (rc) return rc;
  const long tasks = (long)a.W * a.L;
  const long waves = std::min(tasks, (long)ws->cus * 4 * w32::REV32_WAVES_PER_SIMD);
  b.nwaves = (int)waves;
  e = prof_begin(ws, 4, stream);
  w16::fill_int_kernel<<<1, 1, 0, stream>>>(b.queue, 0);
  w32::layer32_adjoint_kernel<<<(unsigned)waves, 64, 0, stream>>>(a, b);
  prof_end(ws, e, stream);
  e = prof_begin(ws, 6, stream);
  w32::start32_adjoint_kernel<<<(unsigned)tasks, 64, 0, stream>>>(a, b);
  prof_end(ws, e, stream);
  e = prof_begin(ws, 5, stream);
  w32::column32_adjoint_kernel<<<(unsigned)((a.W + 255) / 256), 256, 0, stream>>>(a, b);
  if (shared_moments) {
    const long n = (long)a.L * a.M;
    w16::moments16_partial_kernel<<<dim3((unsigned)a.L, w16::MOM_SLICES), 256, 0, stream>>>(
        a, b, ws->mom_part);
    w16::moments16_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(a, b,
                                                                                ws->mom_part);
  }
  prof_end(ws, e, stream);
  return hipGetLastError() == hipSuccess ? ADRT_OK : ADRT_ERR_HIP;
}

int check_solve_args(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                     const adrt_device_outputs* out, adrt_precision precision,
                     adrt_workspace* ws) {
  if (!in->delta_tau) return ADRT_ERR_DELTA_TAU;
  if (!in->single_scat_albedo) return ADRT_ERR_SSA;
  if (!in->phase_moments) return ADRT_ERR_MOMENTS;
  if (cfg->use_diffusion_lower_bc && !in->planck_levels) return ADRT_ERR_DIFFUSION_BC;
  if (!out->flux_up || !out->flux_down || !out->mean_intensity || !out->flux_direct)
    return ADRT_ERR_OUTPUTS;
  if (!ws) return ADRT_ERR_WORKSPACE;
  if (ws->L != cfg->num_layers || ws->n_real != cfg->num_quadrature || ws->M != cfg->num_moments)
    return ADRT_ERR_WORKSPACE;
  // the mixed form runs at quadrature orders 16 (native 9..16, or padded on a
  // reverse-pass workspace) and 32; the mixed-reverse form's forward is the
  // FP64 one
  if (precision != ADRT_PRECISION_F64 && precision != ADRT_PRECISION_MIXED_REVERSE &&
      !(precision == ADRT_PRECISION_MIXED && (ws->Np == 16 || ws->Np == 32)))
    return ADRT_ERR_PRECISION;
  return ADRT_OK;
}

}  // namespace

extern "C" {

int adrt_solve_batch(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                     const adrt_device_outputs* out, adrt_precision precision,
                     adrt_workspace* ws, hipStream_t stream) {
  int rc = check_config(cfg);
  if (rc) return rc;
  if (!in || !out) return ADRT_ERR_NULL;
  if (cfg->num_wavenumbers == 0) return ADRT_OK;  // empty batch: nothing to read or write
  rc = check_solve_args(cfg, in, out, precision, ws);
  if (rc) return rc;
  const bool has_solar = cfg->solar_flux > 0.0 && cfg->solar_mu > 0.0;
  rc = upload_tables(ws, has_solar ? cfg->solar_mu : 1.0, stream);
  if (rc) return rc;
  adrt::dev::FwdArgs a = make_args(cfg, in, out, ws, has_solar);
  ws->kept_valid = false;
  for (int w0 = 0; w0 < cfg->num_wavenumbers; w0 += ws->chunk) {
    a.w0 = w0;
    a.W = std::min(ws->chunk, cfg->num_wavenumbers - w0);
    rc = forward_chunk(a, ws, stream, precision == ADRT_PRECISION_MIXED);
    if (rc) return rc;
  }
  ws->generation = ++g_forward_tokens;
  ws->kept_valid = ws->Iv != nullptr && cfg->num_wavenumbers <= ws->chunk;
  ws->kept_mixed = precision == ADRT_PRECISION_MIXED;  // FP32 records (both other forms: FP64)
  return ADRT_OK;
}

int adrt_source_jacobian_batch(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                               const adrt_device_jacobians* jac, adrt_workspace* ws,
                               hipStream_t stream) {
  int rc = check_config(cfg);
  if (rc) return rc;
  if (!in || !jac) return ADRT_ERR_NULL;
  if (cfg->num_wavenumbers == 0) return ADRT_OK;
  if (!in->delta_tau) return ADRT_ERR_DELTA_TAU;
  if (!in->single_scat_albedo) return ADRT_ERR_SSA;
  if (!in->phase_moments) return ADRT_ERR_MOMENTS;
  if (!jac->flux_up || !jac->flux_down || !jac->mean_intensity) return ADRT_ERR_OUTPUTS;
  if (!ws || !ws->basis || ws->L != cfg->num_layers || ws->n_real != cfg->num_quadrature ||
      ws->M != cfg->num_moments)
    return ADRT_ERR_WORKSPACE;
  // the records need no beam and no sources: R, T and the thermal basis only
  adrt_batch_config c = *cfg;
  c.solar_flux = 0.0;
  rc = upload_tables(ws, 1.0, stream);
  if (rc) return rc;
  adrt_device_inputs ri = *in;
  ri.planck_levels = nullptr;
  ri.surface_emission = nullptr;
  ri.top_emission = nullptr;
  adrt_device_outputs none{nullptr, nullptr, nullptr, nullptr, nullptr};
  adrt::dev::FwdArgs a = make_args(&c, &ri, &none, ws, false);
  a.basis = ws->basis;
  ws->kept_valid = false;
  const int L = cfg->num_layers, U = L + 3;
  for (int w0 = 0; w0 < cfg->num_wavenumbers; w0 += ws->chunk) {
    a.w0 = w0;
    a.W = std::min(ws->chunk, cfg->num_wavenumbers - w0);
    rc = layers_dispatch(a, ws, stream);
    if (rc) return rc;
    const long P = (long)a.W * U;
    JacArgs j{a.W, L, ws->Np, a.rev, ws->lay, ws->basis, ws->lflag, ws->tau_s, ws->tau_cum,
              in->single_scat_albedo, (long)w0, ws->jlay, ws->jlflag, ws->ucol, ws->jucol,
              ws->jtau_s, ws->jtau_cum,
              ws->jB, ws->jsurf, ws->jtop, jac->flux_divergence ? ws->jssa : nullptr};
    jac_records_kernel<<<(unsigned)(P * L), 256, 0, stream>>>(j);
    jac_columns_kernel<<<(unsigned)((P + 127) / 128), 128, 0, stream>>>(j, P);
    if (hipGetLastError() != hipSuccess) return ADRT_ERR_HIP;
    // the adding sweeps and interfaces of the pseudo-columns, outputs into the
    // Jacobians ([W][L+3][L+1]: interface index fastest)
    adrt::dev::FwdArgs p = a;
    p.W = (int)P;
    p.w0 = 0;
    p.has_solar = 0;
    p.B = ws->jB;
    p.surf_w = ws->jsurf;
    p.top_w = ws->jtop;
    p.ssa = ws->jssa;
    p.lay = ws->jlay;
    p.rb = ws->jrb;
    p.lflag = ws->jlflag;
    p.ucol = ws->jucol;
    p.ulist = p.ucount = nullptr;  // the pseudo-columns' flags: no list
    p.tau_s = ws->jtau_s;
    p.tau_cum = ws->jtau_cum;
    p.basis = nullptr;
    p.Iv = p.Av = nullptr;
    const long o = (long)w0 * U * (L + 1);
    p.fu = jac->flux_up + o;
    p.fd = jac->flux_down + o;
    p.mi = jac->mean_intensity + o;
    p.fdir = ws->jfdir;
    p.fdiv = jac->flux_divergence ? jac->flux_divergence + o : nullptr;
    rc = adding_dispatch(p, ws, stream);
    if (rc) return rc;
  }
  return ADRT_OK;
}

int adrt_solve_batch_from_coefficients(const adrt_batch_config* cfg_in,
                                       const adrt_device_coefficients* coef,
                                       double surface_temperature, double top_temperature,
                                       const adrt_device_outputs* out, adrt_precision precision,
                                       adrt_workspace* ws, hipStream_t stream) {
  if (!cfg_in || !coef || !out) return ADRT_ERR_NULL;
  adrt_batch_config cfg = *cfg_in;
  cfg.index_from_bottom = 0;
  cfg.surface_emission = 0.0;
  cfg.top_emission = 0.0;
  int rc = check_config(&cfg);
  if (rc) return rc;
  const int W = cfg.num_wavenumbers, L = cfg.num_layers, M = cfg.num_moments;
  if (W == 0) return ADRT_OK;
  if (!coef->absorption_coeff || !coef->scattering_coeff || !coef->altitude ||
      !coef->temperature || !coef->wavenumber || !coef->phase_moments)
    return ADRT_ERR_NULL;
  if (!ws || ws->L != L) return ADRT_ERR_WORKSPACE;
  const long C = ws->chunk;
  double* tau = ws->raw;
  double* ssa = tau + C * L;
  double* B = ssa + C * L;
  double* Bs = B + C * (L + 1);
  double* Bt = Bs + C;
  for (int w0 = 0; w0 < W; w0 += ws->chunk) {
    const int Wc = std::min(ws->chunk, W - w0);
    raw_optics_kernel<<<(unsigned)((Wc + 127) / 128), 128, 0, stream>>>(
        Wc, w0, W, L, coef->absorption_coeff, coef->scattering_coeff, coef->altitude,
        coef->temperature, coef->wavenumber, coef->cloud_optical_depth, surface_temperature,
        top_temperature, tau, ssa, B, Bs, Bt);
    if (hipGetLastError() != hipSuccess) return ADRT_ERR_HIP;
    adrt_batch_config cc = cfg;
    cc.num_wavenumbers = Wc;
    const double* mom = coef->phase_moments_shared
                            ? coef->phase_moments
                            : coef->phase_moments + (long)w0 * L * M;
    adrt_device_inputs in{tau, ssa, mom, coef->phase_moments_shared, B, Bs, Bt};
    const long o = (long)w0 * (L + 1);
    adrt_device_outputs oc{out->flux_up ? out->flux_up + o : nullptr,
                           out->flux_down ? out->flux_down + o : nullptr,
                           out->mean_intensity ? out->mean_intensity + o : nullptr,
                           out->flux_direct ? out->flux_direct + o : nullptr,
                           out->flux_divergence ? out->flux_divergence + o : nullptr};
    rc = adrt_solve_batch(&cc, &in, &oc, precision, ws, stream);
    if (rc) return rc;
  }
  return ADRT_OK;
}

int adrt_raw_optics(int num_wavenumbers, int num_layers, const adrt_device_coefficients* coef,
                    double surface_temperature, double top_temperature, double* delta_tau,
                    double* single_scat_albedo, double* planck_levels, double* surface_emission,
                    double* top_emission, hipStream_t stream) {
  if (!coef) return ADRT_ERR_NULL;
  if (num_wavenumbers < 0) return ADRT_ERR_NUM_WAVENUMBERS;
  if (num_layers <= 0) return ADRT_ERR_NUM_LAYERS;
  if (num_wavenumbers == 0) return ADRT_OK;
  if (!coef->absorption_coeff || !coef->scattering_coeff || !coef->altitude ||
      !coef->temperature || !coef->wavenumber || !delta_tau || !single_scat_albedo ||
      !planck_levels || !surface_emission || !top_emission)
    return ADRT_ERR_NULL;
  const int W = num_wavenumbers;
  raw_optics_kernel<<<(unsigned)((W + 127) / 128), 128, 0, stream>>>(
      W, 0, W, num_layers, coef->absorption_coeff, coef->scattering_coeff, coef->altitude,
      coef->temperature, coef->wavenumber, coef->cloud_optical_depth, surface_temperature,
      top_temperature, delta_tau, single_scat_albedo, planck_levels, surface_emission,
      top_emission);
  return hipGetLastError() == hipSuccess ? ADRT_OK : ADRT_ERR_HIP;
}

int adrt_raw_optics_adjoint(int num_wavenumbers, int num_layers,
                            const adrt_device_coefficients* coef, double surface_temperature,
                            double top_temperature, const adrt_raw_optics_cotangents* cot,
                            const adrt_raw_gradients* grad, hipStream_t stream) {
  if (!coef || !cot || !grad) return ADRT_ERR_NULL;
  if (num_wavenumbers < 0) return ADRT_ERR_NUM_WAVENUMBERS;
  if (num_layers <= 0) return ADRT_ERR_NUM_LAYERS;
  if (!coef->absorption_coeff || !coef->scattering_coeff || !coef->altitude ||
      !coef->temperature || !coef->wavenumber)
    return ADRT_ERR_NULL;
  const int W = num_wavenumbers, L = num_layers;
  if (W > 0 && (grad->absorption_coeff || grad->scattering_coeff || grad->cloud_optical_depth)) {
    raw_optics_adjoint_kernel<<<(unsigned)((W + 127) / 128), 128, 0, stream>>>(
        W, L, coef->absorption_coeff, coef->scattering_coeff, coef->altitude,
        coef->cloud_optical_depth, cot->delta_tau, cot->single_scat_albedo,
        grad->absorption_coeff, grad->scattering_coeff,
        coef->cloud_optical_depth ? grad->cloud_optical_depth : nullptr);
    if (hipGetLastError() != hipSuccess) return ADRT_ERR_HIP;
  }
  if (grad->temperature || grad->surface_temperature || grad->top_temperature) {
    raw_temperature_adjoint_kernel<<<(unsigned)(L + 3), 256, 0, stream>>>(
        W, L, coef->temperature, coef->wavenumber, surface_temperature, top_temperature,
        cot->planck_levels, cot->surface_emission, cot->top_emission, grad->temperature,
        grad->surface_temperature, grad->top_temperature);
    if (hipGetLastError() != hipSuccess) return ADRT_ERR_HIP;
  }
  return ADRT_OK;
}

int adrt_workspace_generation(const adrt_workspace* ws, unsigned long long* generation) {
  if (!ws || !generation) return ADRT_ERR_NULL;
  *generation = ws->generation;
  return ADRT_OK;
}

int adrt_adjoint_batch(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                       const adrt_device_outputs* out, const adrt_device_cotangents* cot,
                       const adrt_device_gradients* grad, adrt_precision precision,
                       unsigned long long forward_generation, adrt_workspace* ws,
                       hipStream_t stream) {
  using namespace adrt::dev;
  int rc = check_config(cfg);
  if (rc) return rc;
  if (!in || !out || !cot || !grad) return ADRT_ERR_NULL;
  if (cfg->num_wavenumbers == 0) return ADRT_OK;
  rc = check_solve_args(cfg, in, out, precision, ws);
  if (rc) return rc;
  if (!ws->Iv || (ws->Np != 16 && ws->Np != 32)) return ADRT_ERR_ADJOINT;
  // above order 16 the reverse pass runs in FP64 only
  if (ws->Np == 32 && precision != ADRT_PRECISION_F64) return ADRT_ERR_PRECISION;
  const bool has_solar = cfg->solar_flux > 0.0 && cfg->solar_mu > 0.0;
  rc = upload_tables(ws, has_solar ? cfg->solar_mu : 1.0, stream);
  if (rc) return rc;
  FwdArgs a = make_args(cfg, in, out, ws, has_solar);
  a.cfu = cot->flux_up;
  a.cfd = cot->flux_down;
  a.cJ = cot->mean_intensity;
  a.cdir = cot->flux_direct;
  a.cdiv = cot->flux_divergence;
  w16::AdjArgs b{};
  b.slab = ws->slab;
  b.nwaves = ws->nwaves;
  b.nn_cap = ws->nn_cap;
  b.status = ws->status;
  b.queue = ws->queue;
  b.g_taus = ws->g_taus;
  b.g_omgs = ws->g_omgs;
  b.g_Bt = ws->g_Bt;
  b.g_Bb = ws->g_Bb;
  b.g_Ftop = ws->g_Ftop;
  b.g_f = ws->g_f;
  b.g_chi = ws->g_chi;
  b.scot = ws->scot;
  b.d_tau = grad->delta_tau;
  b.d_ssa = grad->single_scat_albedo;
  b.d_mom = grad->phase_moments;
  b.d_B = grad->planck_levels;
  b.d_surf = grad->surface_emission;
  b.d_top = grad->top_emission;
  b.d_albedo = grad->surface_albedo;
  const bool shared = in->phase_moments_shared != 0;
  if (shared && b.d_mom) {  // accumulated over chunks
    const long n = (long)cfg->num_layers * cfg->num_moments;
    w16::fill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(b.d_mom, n, 0.0);
  }
  // forward records: FP32 doubling (MIXED) or FP64 (F64, MIXED_REVERSE); reverse
  // doubling in FP32 for both reduced-precision forms
  const bool mixed = precision == ADRT_PRECISION_MIXED;
  const bool rev_mixed = precision != ADRT_PRECISION_F64;
  if (precision == ADRT_PRECISION_MIXED_REVERSE && ws->Np != 16) return ADRT_ERR_PRECISION;
  if (mixed != ws->kept_mixed) forward_generation = 0;  // kept records are of the other form
  const bool reuse = forward_generation != 0 && ws->kept_valid &&
                     forward_generation == ws->generation && cfg->num_wavenumbers <= ws->chunk;
  for (int w0 = 0; w0 < cfg->num_wavenumbers; w0 += ws->chunk) {
    a.w0 = w0;
    a.W = std::min(ws->chunk, cfg->num_wavenumbers - w0);
    if (!reuse) {
      rc = forward_chunk(a, ws, stream, mixed);
      if (rc) return rc;
      ws->kept_valid = false;
    }
    if (ws->Np == 32) {
      rc = adjoint_chunk32(a, b, ws, stream, shared && b.d_mom);
      if (rc) return rc;
      continue;
    }
    int e = prof_begin(ws, 3, stream);
    const unsigned g = (unsigned)((a.W + 3) / 4);
    if (mixed)
      rc = fork_join(
          ws, stream,
          [&](hipStream_t s) { w16::adding16_kernel<true, true><<<g, 256, 0, s>>>(a); },
          [&](hipStream_t s) { w16::adding16_composite_kernel<true, true><<<g, 256, 0, s>>>(a); });
    else
      rc = fork_join(
          ws, stream,
          [&](hipStream_t s) { w16::adding16_kernel<true, false><<<g, 256, 0, s>>>(a); },
          [&](hipStream_t s) { w16::adding16_composite_kernel<true, false><<<g, 256, 0, s>>>(a); });
    prof_end(ws, e, stream);
    if (rc) return rc;
    const long tasks = (long)a.W * a.L;
    const long blocks = std::min((tasks + 3) / 4, (long)ws->cus * w16::REV_WAVES_PER_SIMD);
    b.nwaves = (int)(blocks * 4);
    e = prof_begin(ws, 4, stream);
    const unsigned task_blocks = (unsigned)((tasks + 3) / 4);
    if (rev_mixed) {
      // its own persistent grid: MREV_WAVES_PER_SIMD waves per SIMD
      w16::AdjArgs bm = b;
      const long mblocks = std::min((tasks + 3) / 4, (long)ws->cus * w16::MREV_WAVES_PER_SIMD);
      bm.nwaves = (int)(mblocks * 4);
      w16::fill_int_kernel<<<1, 1, 0, stream>>>(b.queue, 0);
      w16::layer16m_adjoint_kernel<<<(unsigned)mblocks, 256, 0, stream>>>(a, bm);
      w16::fill_int_kernel<<<1, 1, 0, stream>>>(b.queue, 0);
      w16::layer16_adjoint_kernel<true><<<(unsigned)blocks, 256, 0, stream>>>(a, b);
      prof_end(ws, e, stream);
      e = prof_begin(ws, 6, stream);
      w16::start16_adjoint_kernel<true><<<task_blocks, 256, 0, stream>>>(a, b);
    } else if (a.pord) {
      // two layers per wave (orders up to 8 padded to 16)
      w16::AdjArgs bp = b;
      const long pairs = (long)a.W * (a.L / 2);
      const long pblocks = std::min((pairs + 3) / 4, (long)ws->cus * w16::PAIR_REV_WAVES_PER_SIMD);
      bp.nwaves = (int)(pblocks * 4);
      if (pblocks > 0) {
        w16::fill_int_kernel<<<1, 1, 0, stream>>>(b.queue, 0);
        w16::layer16p_adjoint_kernel<<<(unsigned)pblocks, 256, 0, stream>>>(a, bp);
      }
      w16::fill_int_kernel<<<1, 1, 0, stream>>>(b.queue, 0);
      w16::layer16_adjoint_kernel<false, true><<<(unsigned)blocks, 256, 0, stream>>>(a, b);
      prof_end(ws, e, stream);
      e = prof_begin(ws, 6, stream);
      w16::start16_adjoint_kernel<false><<<task_blocks, 256, 0, stream>>>(a, b);
    } else {
      w16::fill_int_kernel<<<1, 1, 0, stream>>>(b.queue, 0);
      w16::layer16_adjoint_kernel<false><<<(unsigned)blocks, 256, 0, stream>>>(a, b);
      prof_end(ws, e, stream);
      e = prof_begin(ws, 6, stream);
      w16::start16_adjoint_kernel<false><<<task_blocks, 256, 0, stream>>>(a, b);
    }
    prof_end(ws, e, stream);
    e = prof_begin(ws, 5, stream);
    w16::column16_adjoint_kernel<<<(unsigned)((a.W + 255) / 256), 256, 0, stream>>>(a, b);
    if (shared && b.d_mom) {
      const long n = (long)a.L * a.M;
      double* part = ws->mom_part;
      w16::moments16_partial_kernel<<<dim3((unsigned)a.L, w16::MOM_SLICES), 256, 0, stream>>>(
          a, b, part);
      w16::moments16_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(a, b, part);
    }
    prof_end(ws, e, stream);
    if (hipGetLastError() != hipSuccess) return ADRT_ERR_HIP;
  }
  return ADRT_OK;
}

int adrt_adjoint_status(adrt_workspace* ws, hipStream_t stream) {
  if (!ws) return ADRT_ERR_NULL;
  if (!ws->status) return ADRT_ERR_ADJOINT;
  int h = 0;
  ADRT_HIP_TRY(hipMemcpyAsync(&h, ws->status, sizeof(int), hipMemcpyDeviceToHost, stream));
  ADRT_HIP_TRY(hipStreamSynchronize(stream));
  if (h) {
    adrt::dev::w16::fill_int_kernel<<<1, 1, 0, stream>>>(ws->status, 0);
    ADRT_HIP_TRY(hipStreamSynchronize(stream));
    return ADRT_ERR_DOUBLING_STEPS;
  }
  return ADRT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free forward solve: the scratch of the FP64 forward (quadrature
// tables, per-chunk optics, layer store, base composites, phase records) lives
// in one file-scope device array instead of a workspace.  Each GPU holds its
// own copy for the life of the process; its device address is looked up once
// per device (hipGetSymbolAddress) and only that address is used on the host.
// ---------------------------------------------------------------------------
namespace {

static __device__ __align__(256) double scratch[ADRT_NOALLOC_SCRATCH_BYTES / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (dev), looked up on
// the first call on that device
int scratch_address(int dev, char** out) {
  if (dev < 0 || dev >= kMaxDevices) return ADRT_ERR_HIP;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
      return ADRT_ERR_HIP;
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return ADRT_OK;
}

struct QuadTables {
  double mu[ADRT_MAX_QUADRATURE], wt[ADRT_MAX_QUADRATURE];
  double mu0;
  int n, Np, Lmax;
};

// the tables upload_tables writes, computed on the device from the host's
// nodes and weights: mu, wt padded to Np (pad nodes mu = 1, w = 0, P_l = 0),
// P_l(mu_i) [Lmax][Np] and P_l(mu0) [Lmax], by the recurrence of
// legendre_column with the same operations in the same order (no contraction),
// so the tables are identical bit for bit
__global__ void tables_kernel(QuadTables q, double* tab) {
#pragma clang fp contract(off)
  const int i = threadIdx.x;
  double* mu = tab;
  double* wt = mu + q.Np;
  double* Pl = wt + q.Np;
  double* Pl0 = Pl + (size_t)q.Lmax * q.Np;
  if (i > q.Np) return;
  const bool node = i < q.Np;
  if (node) {
    mu[i] = (i < q.n) ? q.mu[i] : 1.0;
    wt[i] = (i < q.n) ? q.wt[i] : 0.0;
  }
  const bool live = !node || i < q.n;
  const double x = node ? q.mu[i < q.n ? i : 0] : q.mu0;
  double pm2 = 0.0, pm1 = 0.0;
  for (int l = 0; l < q.Lmax; ++l) {
    double p;
    if (l == 0) p = 1.0;
    else if (l == 1) p = x;
    else p = ((2 * l - 1) * x * pm1 - (l - 1) * pm2) / l;
    pm2 = pm1;
    pm1 = p;
    if (node) Pl[(size_t)l * q.Np + i] = live ? p : 0.0;
    else Pl0[l] = p;
  }
}

}  // namespace

extern "C" {

int adrt_solve_batch_noalloc(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                             const adrt_device_outputs* out, hipStream_t stream) {
  int rc = check_config(cfg);
  if (rc) return rc;
  if (!in || !out) return ADRT_ERR_NULL;
  if (cfg->num_wavenumbers == 0) return ADRT_OK;
  // a host-side view of the scratch array with the shape of a forward
  // workspace (no member of it allocates: its vectors stay empty)
  adrt_workspace ws;
  ws.L = cfg->num_layers;
  ws.n_real = cfg->num_quadrature;
  ws.Np = padded_order(cfg->num_quadrature);
  ws.M = cfg->num_moments;
  ws.Lmax = std::max(cfg->num_moments, 2 * cfg->num_quadrature);
  rc = check_solve_args(cfg, in, out, ADRT_PRECISION_F64, &ws);
  if (rc) return rc;
  QuadTables q{};
  gauss_legendre01(ws.n_real, q.mu, q.wt);
  double s = 0.0;
  for (int i = 0; i < ws.n_real; ++i) s += q.mu[i] * q.wt[i];
  ws.xfac = 0.5 / s;  // src/solver.cpp:378-383
  const bool has_solar = cfg->solar_flux > 0.0 && cfg->solar_mu > 0.0;
  q.mu0 = has_solar ? cfg->solar_mu : 1.0;
  q.n = ws.n_real;
  q.Np = ws.Np;
  q.Lmax = ws.Lmax;
  // layout: tables and phase records once, then the chunk arrays
  const size_t L = ws.L, Np = ws.Np;
  const size_t tab_bytes = align256((size_t)(2 * Np + ws.Lmax * Np + ws.Lmax) * 8);
  const size_t phase_bytes =
      align256(Np == (size_t)adrt::dev::w16::N ? L * adrt::dev::w16::PH16 * 8 : 0);
  const size_t per_w = per_wavenumber_bytes(ws.L, ws.Np, ws.M, false);
  const size_t fixed = tab_bytes + phase_bytes + 12 * 256;  // + alignment of the chunk arrays
  if (fixed + per_w > ADRT_NOALLOC_SCRATCH_BYTES) return ADRT_ERR_SCRATCH;
  const size_t C = std::min((ADRT_NOALLOC_SCRATCH_BYTES - fixed) / per_w,
                            (size_t)cfg->num_wavenumbers);
  int dev = 0;
  ADRT_HIP_TRY(hipGetDevice(&dev));
  char* base = nullptr;
  rc = scratch_address(dev, &base);
  if (rc) return rc;
  size_t off = 0;
  auto take = [&](size_t b) {
    char* p = base + off;
    off += align256(b);
    return p;
  };
  ws.chunk = (int)C;
  ws.mu = (double*)take(tab_bytes);
  ws.wt = ws.mu + Np;
  ws.Pl = ws.wt + Np;
  ws.Pl0 = ws.Pl + (size_t)ws.Lmax * Np;
  ws.phase = phase_bytes ? (double*)take(phase_bytes) : nullptr;
  ws.tau_s = (double*)take(C * L * 8);
  ws.omega_s = (double*)take(C * L * 8);
  ws.fdm = (double*)take(C * L * 8);
  ws.tau_cum = (double*)take(C * (L + 1) * 8);
  ws.mode = (int*)take(C * L * 4);
  ws.lflag = (int*)take(C * L * 4);
  ws.redo = (int*)take(C * L * 4);
  ws.ucol = (int*)take(C * 4);
  ws.flag = (int*)take(4);
  ws.lay = (double*)take(C * L * (2 * Np * Np + 2 * Np) * 8);
  ws.rb = (double*)take(C * L * (2 * Np * Np + 4 * Np) * 8);
  if (off > ADRT_NOALLOC_SCRATCH_BYTES) return ADRT_ERR_SCRATCH;  // not reached: sized above
  tables_kernel<<<1, 64, 0, stream>>>(q, ws.mu);
  if (hipGetLastError() != hipSuccess) return ADRT_ERR_HIP;
  adrt::dev::FwdArgs a = make_args(cfg, in, out, &ws, has_solar);
  for (int w0 = 0; w0 < cfg->num_wavenumbers; w0 += ws.chunk) {
    a.w0 = w0;
    a.W = std::min(ws.chunk, cfg->num_wavenumbers - w0);
    rc = forward_chunk(a, &ws, stream, false);
    if (rc) return rc;
  }
  return ADRT_OK;
}

int adrt_validate_inputs(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                         adrt_workspace* ws, hipStream_t stream) {
  int rc = check_config(cfg);
  if (rc) return rc;
  if (!in || !in->delta_tau || !in->single_scat_albedo) return ADRT_ERR_NULL;
  if (!ws) return ADRT_ERR_WORKSPACE;
  const long n = (long)cfg->num_wavenumbers * cfg->num_layers;
  if (n == 0) return ADRT_OK;
  ADRT_HIP_TRY(hipMemsetAsync(ws->flag, 0, sizeof(int), stream));
  validate_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(in->delta_tau,
                                                                   in->single_scat_albedo, n,
                                                                   ws->flag);
  int h = 0;
  ADRT_HIP_TRY(hipMemcpyAsync(&h, ws->flag, sizeof(int), hipMemcpyDeviceToHost, stream));
  ADRT_HIP_TRY(hipStreamSynchronize(stream));
  if (h & 1) return ADRT_ERR_NEGATIVE_TAU;
  if (h & 2) return ADRT_ERR_SSA_RANGE;
  return ADRT_OK;
}

int adrt_planck_levels(const double* wn_lo, const double* wn_hi, const double* temperature,
                       int temperature_shared, int num_wavenumbers, int num_levels,
                       int derivative, double* out, adrt_workspace* ws, hipStream_t stream) {
  if (!wn_lo || !wn_hi || !temperature || !out || !ws) return ADRT_ERR_NULL;
  if (num_wavenumbers < 0 || num_levels < 0) return ADRT_ERR_NUM_WAVENUMBERS;
  const long n = (long)num_wavenumbers * num_levels;
  if (n == 0) return ADRT_OK;
  ADRT_HIP_TRY(hipMemsetAsync(ws->flag, 0, sizeof(int), stream));
  planck_levels_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(
      wn_lo, wn_hi, temperature, temperature_shared, num_wavenumbers, num_levels, derivative, out,
      ws->flag);
  int h = 0;
  ADRT_HIP_TRY(hipMemcpyAsync(&h, ws->flag, sizeof(int), hipMemcpyDeviceToHost, stream));
  ADRT_HIP_TRY(hipStreamSynchronize(stream));
  return h ? ADRT_ERR_PLANCK_ARGS : ADRT_OK;
}

int adrt_profile(adrt_workspace* ws, int enable) {
  if (!ws) return ADRT_ERR_WORKSPACE;
  ws->profiling = enable != 0;
  ws->ev_used = 0;
  return ADRT_OK;
}

int adrt_profile_read(adrt_workspace* ws, double* ms, int* launches) {
  if (!ws || !ms) return ADRT_ERR_NULL;
  for (int k = 0; k < ADRT_PROFILE_KINDS; ++k) {
    ms[k] = 0.0;
    if (launches) launches[k] = 0;
  }
  for (size_t k = 0; k < ws->ev_used; ++k) {
    ADRT_HIP_TRY(hipEventSynchronize(ws->ev_pool[2 * k + 1]));
    float t = 0.f;
    ADRT_HIP_TRY(hipEventElapsedTime(&t, ws->ev_pool[2 * k], ws->ev_pool[2 * k + 1]));
    const int kind = ws->ev_kind[k];
    ms[kind] += t;
    if (launches) launches[kind] += 1;
  }
  ws->ev_used = 0;
  return ADRT_OK;
}

int adrt_gauss_legendre(int n, double* nodes, double* weights) {
  if (n < 1 || !nodes || !weights) return ADRT_ERR_NULL;
  gauss_legendre01(n, nodes, weights);
  return ADRT_OK;
}

int adrt_planck(double lo, double hi, double T, double* out) {
  if (!out) return ADRT_ERR_NULL;
  return planck_band(lo, hi, T, out);
}

int adrt_planck_deriv(double lo, double hi, double T, double* out) {
  if (!out) return ADRT_ERR_NULL;
  return planck_band_deriv(lo, hi, T, out);
}

const char* adrt_status_string(int s) {
  switch (s) {
    case ADRT_OK: return "ok";
    case ADRT_ERR_NUM_LAYERS: return "ADConfig: num_layers must be > 0";
    case ADRT_ERR_NUM_QUADRATURE: return "ADConfig: num_quadrature must be >= 2";
    case ADRT_ERR_DELTA_TAU: return "ADConfig: delta_tau size mismatch";
    case ADRT_ERR_SSA: return "ADConfig: single_scat_albedo size mismatch";
    case ADRT_ERR_MOMENTS: return "ADConfig: phase_function_moments size mismatch";
    case ADRT_ERR_NEGATIVE_TAU: return "ADConfig: delta_tau < 0";
    case ADRT_ERR_SSA_RANGE: return "ADConfig: single_scat_albedo out of [0,1]";
    case ADRT_ERR_SURFACE_ALBEDO: return "ADConfig: surface_albedo out of [0,1]";
    case ADRT_ERR_SOLAR_MU: return "ADConfig: solar_mu must be in (0,1] when solar_flux > 0";
    case ADRT_ERR_TEMPERATURE_SIZE: return "ADConfig: temperature size must be num_layers+1";
    case ADRT_ERR_TEMPERATURE: return "ADConfig: temperature < 0";
    case ADRT_ERR_WAVENUMBER_RANGE: return "ADConfig: invalid wavenumber range";
    case ADRT_ERR_SURFACE_TEMPERATURE:
      return "ADConfig: surface_temperature requires use_thermal_emission (use surface_emission "
             "for raw thermal sources)";
    case ADRT_ERR_TOP_TEMPERATURE:
      return "ADConfig: top_temperature requires use_thermal_emission (use top_emission for raw "
             "thermal sources)";
    case ADRT_ERR_DIFFUSION_BC:
      return "ADConfig: use_diffusion_lower_bc requires use_thermal_emission or planck_levels";
    case ADRT_ERR_PLANCK_ARGS: return "planckFunction: invalid arguments";
    case ADRT_ERR_QUADRATURE_MAX: return "num_quadrature exceeds 32";
    case ADRT_ERR_NUM_WAVENUMBERS: return "num_wavenumbers must be >= 0";
    case ADRT_ERR_OUTPUTS: return "required output buffer is NULL";
    case ADRT_ERR_WORKSPACE: return "workspace missing or built for another shape";
    case ADRT_ERR_HIP: return "HIP runtime error";
    case ADRT_ERR_PRECISION: return "unknown or unsupported precision";
    case ADRT_ERR_NULL: return "NULL argument";
    case ADRT_ERR_ADJOINT:
      return "reverse pass needs a workspace created with ADRT_WORKSPACE_ADJOINT (num_quadrature <= 16)";
    case ADRT_ERR_DOUBLING_STEPS:
      return "a layer needs more doubling steps than the reverse pass checkpoints (delta_tau too large)";
    case ADRT_ERR_SCRATCH:
      return "adrt_solve_batch_noalloc: one column of this shape (num_layers, num_quadrature, "
             "num_moments) needs more than the ADRT_NOALLOC_SCRATCH_BYTES of scratch it holds";
    default: return "unknown status";
  }
}

const char* adrt_version(void) { return "adrt-mi355x 0.1"; }

}  // extern "C"
