/*
 * adrt_hip.hpp -- C++ host layer over the C-ABI of adrt_hip.h.
 *
 * The same names, fields and error behaviour as the reference's C++ entry
 * points, running the columns on the GPU through the C-ABI:
 *
 *   adrt::ADConfig, adrt::RTOutput     <- src/adding_doubling.hpp:32-247
 *   adrt::solve(const ADConfig&[, SolverWorkspace&])
 *                                      <- src/adding_doubling.hpp:253-269
 *   adrt::gaussLegendre, planckFunction, planckFunctionDeriv
 *                                      <- src/adding_doubling.hpp:277-296, src/planck.hpp
 *   adrt::hip::BatchConfig, DeviceData, solveBatch, solveBatchHost
 *                                      <- adrt::cuda::* of src_cuda/cuda_solver.cuh:18-98
 *   adrt::hip::RawDeviceData, SolverWorkspaceGPU, solveBatchFromCoefficients
 *                                      <- src_cuda/cuda_solver.cuh:104-179
 *
 * Invalid configurations throw std::invalid_argument with the reference's
 * messages; HIP failures throw std::runtime_error.  Header only: include it
 * from a HIP or C++ translation unit (with the ROCm include path) and link
 * libadrt_hip.so and the HIP runtime (-lamdhip64).
 *
 * Differences from the reference, all deliberate:
 *  - solve() runs the column on the current HIP device in FP64 (a batch of
 *    one).  Temperature Jacobians come from one batched solve of L+3 copies
 *    of the column with unit Planck sources.
 *  - hip::DeviceData is FP64.  Its outputs are TOA values per wavenumber
 *    (flux_direct: at the surface) as in the reference, unless
 *    all_interfaces is set, in which case they are [nwav][nlay+1] and
 *    mean_intensity / flux_divergence are available too.
 *  - Temporaries come from a per-thread cache that grows with the batch, not
 *    from a malloc/free per call.
 */
#ifndef ADRT_HIP_HPP
#define ADRT_HIP_HPP

#include "adrt_hip.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace adrt {

namespace detail {

inline void check(int status) {
  if (status == ADRT_OK) return;
  if (status == ADRT_ERR_HIP) throw std::runtime_error(adrt_status_string(status));
  throw std::invalid_argument(adrt_status_string(status));
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer of doubles that only grows.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
  }
  double* get(std::size_t n) {
    if (n > n_) {
      if (p_) (void)hipFree(p_);
      p_ = nullptr;
      n_ = 0;
      hip_check(hipMalloc((void**)&p_, std::max<std::size_t>(n, 1) * sizeof(double)), "hipMalloc");
      n_ = n;
    }
    return p_;
  }
  double* data() const { return p_; }

 private:
  double* p_ = nullptr;
  std::size_t n_ = 0;
};

// adrt_workspace kept for one column shape and flag set; recreated on change.
class Workspace {
 public:
  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace() { reset(); }
  adrt_workspace* get(const adrt_batch_config& c, int flags) {
    if (ws_ && c.num_layers == L_ && c.num_quadrature == N_ && c.num_moments == M_ &&
        flags == flags_ && c.num_wavenumbers <= W_)
      return ws_;
    reset();
    check(adrt_workspace_create_ex(&c, 0, flags, &ws_));
    L_ = c.num_layers;
    N_ = c.num_quadrature;
    M_ = c.num_moments;
    W_ = c.num_wavenumbers;
    flags_ = flags;
    return ws_;
  }
  void reset() {
    if (ws_) adrt_workspace_destroy(ws_);
    ws_ = nullptr;
  }

 private:
  adrt_workspace* ws_ = nullptr;
  int L_ = 0, N_ = 0, M_ = 0, W_ = 0, flags_ = -1;
};

inline void upload(double* dst, const std::vector<double>& src, hipStream_t s) {
  if (!src.empty())
    hip_check(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice,
                             s),
              "hipMemcpyAsync");
}

inline void download(std::vector<double>& dst, const double* src, std::size_t n, hipStream_t s) {
  dst.assign(n, 0.0);
  if (n)
    hip_check(hipMemcpyAsync(dst.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
}

}  // namespace detail

// ============================================================================
//  Column interface (src/adding_doubling.hpp)
// ============================================================================

/// Results at every interface, index 0 = TOA, index num_layers = bottom.
struct RTOutput {
  std::vector<double> flux_up;
  std::vector<double> flux_down;
  std::vector<double> mean_intensity;   ///< includes the direct beam F_dir / (4 pi mu0)
  std::vector<double> flux_divergence;  ///< 4 pi (1 - omega) (J - B), layer above each level
  std::vector<double> flux_direct;
  /// [interface][dof], dof 0..num_layers = level temperatures, num_layers+1 =
  /// surface temperature; d/dB when driven by planck_levels.  Filled only with
  /// ADConfig::compute_temperature_jacobian.
  std::vector<std::vector<double>> flux_up_temperature_jac;
  std::vector<std::vector<double>> flux_down_temperature_jac;
  std::vector<std::vector<double>> mean_intensity_temperature_jac;
  std::vector<std::vector<double>> flux_divergence_temperature_jac;
};

/// Column configuration, fields and defaults as the reference's.
class ADConfig {
 public:
  int num_layers = 0;
  int num_quadrature = 8;

  bool use_thermal_emission = false;
  bool use_delta_m = false;
  bool use_diffusion_lower_bc = false;
  bool index_from_bottom = false;
  bool compute_temperature_jacobian = false;

  double surface_albedo = 0.0;
  double surface_emission = 0.0;
  double surface_temperature = -1.0;
  double top_emission = 0.0;
  double top_temperature = -1.0;
  double solar_flux = 0.0;
  double solar_mu = 1.0;

  double wavenumber_low = 0.0;
  double wavenumber_high = 0.0;

  std::vector<double> delta_tau;
  std::vector<double> single_scat_albedo;
  std::vector<double> temperature;
  std::vector<double> planck_levels;
  std::vector<std::vector<double>> phase_function_moments;

  ADConfig() = default;
  ADConfig(int nlyr, int nquad = 8) : num_layers(nlyr), num_quadrature(nquad) {}

  /// Arrays sized for the current dimensions and flags: 2N moments (2N + 1
  /// with delta-M) per layer with chi_0 = 1, temperatures when thermal.
  void allocate() {
    delta_tau.assign(num_layers, 0.0);
    single_scat_albedo.assign(num_layers, 0.0);
    const int nmom = use_delta_m ? 2 * num_quadrature + 1 : 2 * num_quadrature;
    phase_function_moments.assign(num_layers, std::vector<double>(nmom, 0.0));
    for (auto& pm : phase_function_moments) pm[0] = 1.0;
    if (use_thermal_emission) temperature.assign(num_layers + 1, 0.0);
  }

  /// The reference's checks, in its order and with its messages.
  void validate() const {
    auto fail = [](const std::string& m) { throw std::invalid_argument("ADConfig: " + m); };
    if (num_layers <= 0) fail("num_layers must be > 0");
    if (num_quadrature < 2) fail("num_quadrature must be >= 2");
    if ((int)delta_tau.size() != num_layers) fail("delta_tau size mismatch");
    if ((int)single_scat_albedo.size() != num_layers) fail("single_scat_albedo size mismatch");
    if ((int)phase_function_moments.size() != num_layers)
      fail("phase_function_moments size mismatch");
    for (int l = 0; l < num_layers; ++l) {
      if (delta_tau[l] < 0.0) fail("delta_tau[" + std::to_string(l) + "] < 0");
      if (single_scat_albedo[l] < 0.0 || single_scat_albedo[l] > 1.0)
        fail("single_scat_albedo[" + std::to_string(l) + "] out of [0,1]");
    }
    if (surface_albedo < 0.0 || surface_albedo > 1.0) fail("surface_albedo out of [0,1]");
    if (solar_flux > 0.0 && (solar_mu <= 0.0 || solar_mu > 1.0))
      fail("solar_mu must be in (0,1] when solar_flux > 0");
    if (use_thermal_emission) {
      if ((int)temperature.size() != num_layers + 1)
        fail("temperature size must be num_layers+1");
      for (int l = 0; l <= num_layers; ++l)
        if (temperature[l] < 0.0) fail("temperature[" + std::to_string(l) + "] < 0");
      if (wavenumber_low < 0.0 || wavenumber_high < wavenumber_low)
        fail("invalid wavenumber range");
    } else if (surface_temperature >= 0.0) {
      fail("surface_temperature requires use_thermal_emission "
           "(use surface_emission for raw thermal sources)");
    } else if (top_temperature >= 0.0) {
      fail("top_temperature requires use_thermal_emission "
           "(use top_emission for raw thermal sources)");
    }
    if (use_diffusion_lower_bc && !use_thermal_emission &&
        (int)planck_levels.size() != num_layers + 1)
      fail("use_diffusion_lower_bc requires use_thermal_emission or planck_levels");
  }

  /// Henyey-Greenstein moments g^k (layer lc, or all for lc < 0).
  void setHenyeyGreenstein(double g, int lc = -1) {
    for (int l = first(lc); l < last(lc); ++l) {
      double gk = 1.0;
      for (double& c : phase_function_moments[l]) {
        c = gk;
        gk *= g;
      }
    }
  }
  /// f g1^k + (1 - f) g2^k.
  void setDoubleHenyeyGreenstein(double f, double g1, double g2, int lc = -1) {
    for (int l = first(lc); l < last(lc); ++l) {
      double a = 1.0, b = 1.0;
      for (double& c : phase_function_moments[l]) {
        c = f * a + (1.0 - f) * b;
        a *= g1;
        b *= g2;
      }
    }
  }
  void setIsotropic(int lc = -1) {
    for (int l = first(lc); l < last(lc); ++l) {
      std::fill(phase_function_moments[l].begin(), phase_function_moments[l].end(), 0.0);
      phase_function_moments[l][0] = 1.0;
    }
  }
  /// chi = (1, 0, 0.1, 0, ...).
  void setRayleigh(int lc = -1) {
    for (int l = first(lc); l < last(lc); ++l) {
      std::fill(phase_function_moments[l].begin(), phase_function_moments[l].end(), 0.0);
      phase_function_moments[l][0] = 1.0;
      if (phase_function_moments[l].size() > 2) phase_function_moments[l][2] = 0.1;
    }
  }

 private:
  int first(int lc) const { return lc < 0 ? 0 : lc; }
  int last(int lc) const { return lc < 0 ? num_layers : lc + 1; }
};

inline void gaussLegendre(int n, std::vector<double>& nodes, std::vector<double>& weights) {
  nodes.assign(std::max(n, 0), 0.0);
  weights.assign(std::max(n, 0), 0.0);
  detail::check(adrt_gauss_legendre(n, nodes.data(), weights.data()));
}

inline double planckFunction(double wnumlo, double wnumhi, double temp) {
  double b = 0.0;
  detail::check(adrt_planck(wnumlo, wnumhi, temp, &b));
  return b;
}

inline double planckFunctionDeriv(double wnumlo, double wnumhi, double temp) {
  double b = 0.0;
  detail::check(adrt_planck_deriv(wnumlo, wnumhi, temp, &b));
  return b;
}

/// Device memory and workspaces reused across solve() calls of one thread
/// (the reference's SolverWorkspace, src/workspace.hpp:5-48).
class SolverWorkspace {
 public:
  SolverWorkspace() = default;
  SolverWorkspace(const SolverWorkspace&) = delete;
  SolverWorkspace& operator=(const SolverWorkspace&) = delete;
  ~SolverWorkspace() {
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  hipStream_t stream() {
    if (!stream_) detail::hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
    return stream_;
  }
  detail::Workspace forward, jacobian;
  detail::DeviceBuffer in[6], out[5];

 private:
  hipStream_t stream_ = nullptr;
};

namespace detail {

// Host-side column arrays in the C-ABI's form (user indexing kept: the
// library applies index_from_bottom itself).
struct Column {
  adrt_batch_config cfg{};
  std::vector<double> tau, ssa, mom, B;
  bool have_B = false;
  double Bs = 0.0, Bt = 0.0;
  bool surface_independent = false;  // surface emission is its own Jacobian dof
};

inline Column column_of(const ADConfig& c) {
  Column k;
  const int L = c.num_layers;
  std::size_t M = 1;
  for (const auto& pm : c.phase_function_moments) M = std::max(M, pm.size());
  k.cfg.num_wavenumbers = 1;
  k.cfg.num_layers = L;
  k.cfg.num_quadrature = c.num_quadrature;
  k.cfg.num_moments = (int)M;
  k.cfg.use_delta_m = c.use_delta_m;
  k.cfg.use_diffusion_lower_bc = c.use_diffusion_lower_bc;
  k.cfg.index_from_bottom = c.index_from_bottom;
  k.cfg.surface_albedo = c.surface_albedo;
  k.cfg.solar_flux = c.solar_flux;
  k.cfg.solar_mu = c.solar_mu;
  k.cfg.spectrum_scaling = 1.0;
  k.tau = c.delta_tau;
  k.ssa = c.single_scat_albedo;
  // shorter moment lists are padded with zeros, which adds nothing to the
  // Legendre sums and leaves the delta-M fraction at 0 (no scaling) as a
  // list of 2N moments has it
  k.mom.assign((std::size_t)L * M, 0.0);
  for (int l = 0; l < L; ++l)
    std::copy(c.phase_function_moments[l].begin(), c.phase_function_moments[l].end(),
              k.mom.begin() + (std::size_t)l * M);
  // Planck values and boundary emission (src/solver.cpp:385-419)
  const int top = c.index_from_bottom ? L : 0, bottom = c.index_from_bottom ? 0 : L;
  if (c.use_thermal_emission) {
    k.B.resize(L + 1);
    for (int l = 0; l <= L; ++l)
      k.B[l] = planckFunction(c.wavenumber_low, c.wavenumber_high, c.temperature[l]);
    k.have_B = true;
    k.Bs = c.surface_temperature >= 0.0
               ? planckFunction(c.wavenumber_low, c.wavenumber_high, c.surface_temperature)
               : k.B[bottom];
    k.Bt = c.top_temperature >= 0.0
               ? planckFunction(c.wavenumber_low, c.wavenumber_high, c.top_temperature)
               : k.B[top];
  } else {
    if ((int)c.planck_levels.size() == L + 1) {
      k.B = c.planck_levels;
      k.have_B = true;
    }
    k.Bs = c.surface_emission;
    k.Bt = c.top_emission;
  }
  k.cfg.surface_emission = k.Bs;
  k.cfg.top_emission = k.Bt;
  const bool has_surface =
      !c.use_diffusion_lower_bc && (c.surface_albedo > 0.0 || k.Bs > 0.0);
  k.surface_independent =
      has_surface && (c.use_thermal_emission ? c.surface_temperature >= 0.0 : k.have_B);
  return k;
}

}  // namespace detail

/// Temperature (or Planck-level) Jacobians, as computeTemperatureJacobian
/// (src/solver.cpp:117-341).  R and T do not depend on temperature and every
/// output is affine in the Planck values, so the derivative of the outputs by
/// level value B_m is the solve with B = e_m and no beam (the reference runs
/// the same linear recursion in forward mode).  adrt_source_jacobian_batch
/// runs the column's doubling once and the adding and interface solves for
/// the L+1 level sources and the surface and top emission.  Any quadrature
/// order.
inline void temperature_jacobians(const ADConfig& c, const detail::Column& k, RTOutput& r,
                                  SolverWorkspace& ws) {
  const int L = c.num_layers, nif = L + 1, ndof = L + 2;
  const int top = c.index_from_bottom ? L : 0, bottom = c.index_from_bottom ? 0 : L;
  adrt_batch_config cfg = k.cfg;
  cfg.num_wavenumbers = 1;
  cfg.spectrum_scaling = 1.0;
  hipStream_t s = ws.stream();
  adrt_workspace* jw = ws.jacobian.get(cfg, ADRT_WORKSPACE_JACOBIAN);
  const int U = nif + 2;  // sources: level m <= L, L+1 surface emission, L+2 top emission
  double* dtau = ws.in[0].get((std::size_t)L);
  double* dssa = ws.in[1].get((std::size_t)L);
  double* dmom = ws.in[2].get(k.mom.size());
  detail::upload(dtau, k.tau, s);
  detail::upload(dssa, k.ssa, s);
  detail::upload(dmom, k.mom, s);
  adrt_device_inputs in{dtau, dssa, dmom, 1, nullptr, nullptr, nullptr};
  double* o[4];
  for (int q = 0; q < 4; ++q) o[q] = ws.out[q].get((std::size_t)U * nif);
  adrt_device_jacobians jac{o[0], o[1], o[2], o[3]};
  detail::check(adrt_source_jacobian_batch(&cfg, &in, &jac, jw, s));
  std::vector<double> h[4];
  for (int a = 0; a < 4; ++a) detail::download(h[a], o[a], (std::size_t)U * nif, s);
  detail::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  // d(output a at interface i) / d(B_m): column m, interface i
  auto d = [&](int a, int i, int col) { return h[a][(std::size_t)col * nif + i]; };
  // Planck chain rule (src/solver.cpp:307-327)
  std::vector<double> chain(ndof, 1.0);
  if (c.use_thermal_emission) {
    for (int m = 0; m <= L; ++m)
      chain[m] = planckFunctionDeriv(c.wavenumber_low, c.wavenumber_high, c.temperature[m]);
    chain[L + 1] = k.surface_independent ? planckFunctionDeriv(c.wavenumber_low,
                                                                c.wavenumber_high,
                                                                c.surface_temperature)
                                         : 0.0;
  } else if (!k.surface_independent) {
    chain[L + 1] = 0.0;
  }
  const bool top_dep = c.use_thermal_emission && c.top_temperature < 0.0;
  std::vector<std::vector<double>>* J[4] = {&r.flux_up_temperature_jac,
                                            &r.flux_down_temperature_jac,
                                            &r.mean_intensity_temperature_jac,
                                            &r.flux_divergence_temperature_jac};
  for (int a = 0; a < 4; ++a) {
    J[a]->assign(nif, std::vector<double>(ndof, 0.0));
    for (int i = 0; i < nif; ++i) {
      std::vector<double>& row = (*J[a])[i];
      if (k.have_B)
        for (int m = 0; m <= L; ++m) row[m] = d(a, i, m);
      if (top_dep) row[top] += d(a, i, nif + 1);
      if (k.surface_independent)
        row[L + 1] = d(a, i, nif);
      else if (c.use_thermal_emission)
        row[bottom] += d(a, i, nif);
      for (int m = 0; m < ndof; ++m) row[m] *= chain[m];
    }
  }
}

/// One column on the GPU, FP64.
inline RTOutput solve(const ADConfig& config, SolverWorkspace& ws) {
  config.validate();
  detail::Column k = detail::column_of(config);
  const int L = config.num_layers, nif = L + 1;
  hipStream_t s = ws.stream();
  adrt_workspace* fw = ws.forward.get(k.cfg, 0);
  double* dtau = ws.in[0].get(L);
  double* dssa = ws.in[1].get(L);
  double* dmom = ws.in[2].get(k.mom.size());
  double* dB = k.have_B ? ws.in[3].get(nif) : nullptr;
  detail::upload(dtau, k.tau, s);
  detail::upload(dssa, k.ssa, s);
  detail::upload(dmom, k.mom, s);
  if (dB) detail::upload(dB, k.B, s);
  adrt_device_inputs in{dtau, dssa, dmom, 1, dB, nullptr, nullptr};
  double* o[5];
  for (int q = 0; q < 5; ++q) o[q] = ws.out[q].get(nif);
  adrt_device_outputs out{o[0], o[1], o[2], o[3], o[4]};
  detail::check(adrt_solve_batch(&k.cfg, &in, &out, ADRT_PRECISION_F64, fw, s));
  RTOutput r;
  detail::download(r.flux_up, o[0], nif, s);
  detail::download(r.flux_down, o[1], nif, s);
  detail::download(r.mean_intensity, o[2], nif, s);
  detail::download(r.flux_direct, o[3], nif, s);
  detail::download(r.flux_divergence, o[4], nif, s);
  detail::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (config.compute_temperature_jacobian) temperature_jacobians(config, k, r, ws);
  return r;
}

inline RTOutput solve(const ADConfig& config) {
  thread_local SolverWorkspace ws;
  return solve(config, ws);
}

// ============================================================================
//  Batched interface (src_cuda/cuda_solver.cuh), FP64 device buffers
// ============================================================================
namespace hip {

struct BatchConfig {
  int num_wavenumbers = 0;
  int num_layers = 0;
  int num_quadrature = 8;
  int num_moments_max = 0;

  bool use_delta_m = false;
  bool use_thermal_emission = false;
  bool use_diffusion_lower_bc = false;
  bool index_from_bottom = false;

  double surface_albedo = 0.0;
  double surface_emission = 0.0;
  double surface_temperature = -1.0;
  double top_emission = 0.0;
  double top_temperature = -1.0;
  double solar_flux = 0.0;
  double solar_mu = 1.0;

  double wavenumber_low = 0.0;
  double wavenumber_high = 0.0;

  double spectrum_scaling = 1.0;
  adrt_precision precision = ADRT_PRECISION_F64;
};

/// Caller-owned device memory, array[wav * stride + layer_or_level].
struct DeviceData {
  const double* delta_tau = nullptr;           ///< [nwav * nlay]
  const double* single_scat_albedo = nullptr;  ///< [nwav * nlay]
  const double* phase_moments = nullptr;       ///< [nwav * nlay * nmom] or [nlay * nmom]
  bool phase_moments_shared = false;
  const double* temperature = nullptr;      ///< [nwav * nlev] (use_thermal_emission)
  const double* planck_levels = nullptr;    ///< [nwav * nlev] otherwise, or nullptr
  const double* surface_emission = nullptr;  ///< [nwav] or nullptr (BatchConfig value)
  const double* top_emission = nullptr;      ///< [nwav] or nullptr

  bool all_interfaces = false;  ///< outputs [nwav * nlev] instead of TOA [nwav]
  double* flux_up = nullptr;
  double* flux_down = nullptr;
  double* flux_direct = nullptr;      ///< at the surface (TOA mode) or every interface
  double* mean_intensity = nullptr;   ///< all_interfaces only
  double* flux_divergence = nullptr;  ///< all_interfaces only
};

namespace detail_batch {
struct Cache {
  adrt::detail::Workspace ws;
  adrt::detail::DeviceBuffer band_lo, band_hi, planck, surf, top, out[5];
};
inline Cache& cache() {
  thread_local Cache c;
  return c;
}
}  // namespace detail_batch

/// Solve config.num_wavenumbers columns on the caller's stream.  Throws
/// std::invalid_argument for invalid configurations (the reference returns
/// nothing and checks nothing).
inline void solveBatch(const BatchConfig& config, const DeviceData& data,
                       hipStream_t stream = 0) {
  using adrt::detail::check;
  using adrt::detail::hip_check;
  auto& C = detail_batch::cache();
  const int W = config.num_wavenumbers, L = config.num_layers, nif = L + 1;
  if (W < 0) check(ADRT_ERR_NUM_WAVENUMBERS);
  if (config.use_thermal_emission) {
    if (config.wavenumber_low < 0.0 || config.wavenumber_high < config.wavenumber_low)
      check(ADRT_ERR_WAVENUMBER_RANGE);
    if (!data.temperature) check(ADRT_ERR_TEMPERATURE_SIZE);
  } else if (config.surface_temperature >= 0.0) {
    check(ADRT_ERR_SURFACE_TEMPERATURE);
  } else if (config.top_temperature >= 0.0) {
    check(ADRT_ERR_TOP_TEMPERATURE);
  }
  adrt_batch_config c{};
  c.num_wavenumbers = W;
  c.num_layers = L;
  c.num_quadrature = config.num_quadrature;
  c.num_moments = config.num_moments_max;
  c.use_delta_m = config.use_delta_m;
  c.use_diffusion_lower_bc = config.use_diffusion_lower_bc;
  c.index_from_bottom = config.index_from_bottom;
  c.surface_albedo = config.surface_albedo;
  c.surface_emission = config.surface_emission;
  c.top_emission = config.top_emission;
  c.solar_flux = config.solar_flux;
  c.solar_mu = config.solar_mu;
  c.spectrum_scaling = config.spectrum_scaling;
  if (W == 0) return;
  adrt_workspace* ws =
      C.ws.get(c, config.precision == ADRT_PRECISION_MIXED ? ADRT_WORKSPACE_ADJOINT : 0);
  adrt_device_inputs in{data.delta_tau, data.single_scat_albedo, data.phase_moments,
                        data.phase_moments_shared, data.planck_levels, data.surface_emission,
                        data.top_emission};
  if (config.use_thermal_emission) {
    // band Planck at every level (the reference evaluates it per level in
    // the kernel, src_cuda/cuda_solver.cu:182-186, 241-249, 432-440)
    std::vector<double> lo(W, config.wavenumber_low), hi(W, config.wavenumber_high);
    double* dlo = C.band_lo.get(W);
    double* dhi = C.band_hi.get(W);
    adrt::detail::upload(dlo, lo, stream);
    adrt::detail::upload(dhi, hi, stream);
    double* B = C.planck.get((std::size_t)W * nif);
    check(adrt_planck_levels(dlo, dhi, data.temperature, 0, W, nif, 0, B, ws, stream));
    in.planck_levels = B;
    const int top = config.index_from_bottom ? L : 0;
    const int bottom = config.index_from_bottom ? 0 : L;
    if (config.surface_temperature >= 0.0) {
      c.surface_emission =
          planckFunction(config.wavenumber_low, config.wavenumber_high, config.surface_temperature);
      in.surface_emission = nullptr;
    } else {
      double* Bs = C.surf.get(W);
      hip_check(hipMemcpy2DAsync(Bs, sizeof(double), B + bottom, nif * sizeof(double),
                                 sizeof(double), W, hipMemcpyDeviceToDevice, stream),
                "hipMemcpy2DAsync");
      in.surface_emission = Bs;
    }
    if (config.top_temperature >= 0.0) {
      c.top_emission =
          planckFunction(config.wavenumber_low, config.wavenumber_high, config.top_temperature);
      in.top_emission = nullptr;
    } else {
      double* Bt = C.top.get(W);
      hip_check(hipMemcpy2DAsync(Bt, sizeof(double), B + top, nif * sizeof(double),
                                 sizeof(double), W, hipMemcpyDeviceToDevice, stream),
                "hipMemcpy2DAsync");
      in.top_emission = Bt;
    }
  }
  if (data.all_interfaces) {
    adrt_device_outputs out{data.flux_up, data.flux_down, data.mean_intensity, data.flux_direct,
                            data.flux_divergence};
    double* o[4];
    for (int q = 0; q < 4; ++q) o[q] = C.out[q].get((std::size_t)W * nif);
    if (!out.flux_up) out.flux_up = o[0];
    if (!out.flux_down) out.flux_down = o[1];
    if (!out.mean_intensity) out.mean_intensity = o[2];
    if (!out.flux_direct) out.flux_direct = o[3];
    check(adrt_solve_batch(&c, &in, &out, config.precision, ws, stream));
    return;
  }
  double* o[4];
  for (int q = 0; q < 4; ++q) o[q] = C.out[q].get((std::size_t)W * nif);
  adrt_device_outputs out{o[0], o[1], o[2], o[3], nullptr};
  check(adrt_solve_batch(&c, &in, &out, config.precision, ws, stream));
  // TOA values (flux_direct at the surface) into the caller's [nwav] arrays
  const int toa = config.index_from_bottom ? L : 0;
  const int surf = config.index_from_bottom ? 0 : L;
  auto column = [&](double* dst, const double* src, int level) {
    if (dst)
      hip_check(hipMemcpy2DAsync(dst, sizeof(double), src + level, nif * sizeof(double),
                                 sizeof(double), W, hipMemcpyDeviceToDevice, stream),
                "hipMemcpy2DAsync");
  };
  column(data.flux_up, o[0], toa);
  column(data.flux_down, o[1], toa);
  column(data.flux_direct, o[3], surf);
}

/// Raw retrieval inputs (src_cuda/cuda_solver.cuh:104-134), FP64: coefficient
/// arrays level-major array[level * nwav + wav] with the bottom of the
/// atmosphere at index 0; phase moments TOA first as in DeviceData.
struct RawDeviceData {
  const double* absorption_coeff = nullptr;     ///< [nlev * nwav] (cm^-1)
  const double* scattering_coeff = nullptr;     ///< [nlev * nwav] (cm^-1)
  const double* altitude = nullptr;             ///< [nlev] (cm)
  const double* temperature = nullptr;          ///< [nlev] (K)
  const double* wavenumber = nullptr;           ///< [nwav] (cm^-1)
  const double* phase_moments = nullptr;        ///< [nlay * nmom] or [nwav * nlay * nmom]
  bool phase_moments_shared = false;
  const double* cloud_optical_depth = nullptr;  ///< [nlay * nwav] (BOA first) or nullptr

  bool all_interfaces = false;  ///< outputs [nwav * nlev] (TOA first) instead of TOA [nwav]
  double* flux_up = nullptr;
  double* flux_down = nullptr;
  double* flux_direct = nullptr;      ///< at the surface (TOA mode) or every interface
  double* mean_intensity = nullptr;   ///< all_interfaces only
  double* flux_divergence = nullptr;  ///< all_interfaces only
};

/// Persistent workspace of the raw path (src_cuda/cuda_solver.cuh:140-156):
/// the library workspace (which also holds the per-wavenumber optical depths,
/// albedos and Planck values) and the all-interface outputs of TOA mode.
struct SolverWorkspaceGPU {
  adrt::detail::Workspace ws;
  adrt::detail::DeviceBuffer out[4];
};

/// Solve from raw coefficients (src_cuda/cuda_solver.cuh:164-179).  Optical
/// depths by the trapezoid rule, single-wavenumber Planck values, surface /
/// top emission from config.surface_temperature / top_temperature when >= 0
/// (else the bottom / top level), for every quadrature order alike; the
/// band-Planck fields, use_thermal_emission, surface_emission, top_emission and
/// index_from_bottom of config are not used.  Throws std::invalid_argument for
/// invalid configurations.
inline void solveBatchFromCoefficients(const BatchConfig& config, const RawDeviceData& data,
                                       SolverWorkspaceGPU& w, hipStream_t stream = 0) {
  using adrt::detail::check;
  using adrt::detail::hip_check;
  const int W = config.num_wavenumbers, L = config.num_layers, nif = L + 1;
  if (W < 0) check(ADRT_ERR_NUM_WAVENUMBERS);
  adrt_batch_config c{};
  c.num_wavenumbers = W;
  c.num_layers = L;
  c.num_quadrature = config.num_quadrature;
  c.num_moments = config.num_moments_max;
  c.use_delta_m = config.use_delta_m;
  c.use_diffusion_lower_bc = config.use_diffusion_lower_bc;
  c.surface_albedo = config.surface_albedo;
  c.solar_flux = config.solar_flux;
  c.solar_mu = config.solar_mu;
  c.spectrum_scaling = config.spectrum_scaling;
  if (W == 0) return;
  adrt_workspace* ws =
      w.ws.get(c, config.precision == ADRT_PRECISION_MIXED ? ADRT_WORKSPACE_ADJOINT : 0);
  const adrt_device_coefficients k{data.absorption_coeff, data.scattering_coeff, data.altitude,
                                   data.temperature,      data.wavenumber,        data.phase_moments,
                                   data.phase_moments_shared, data.cloud_optical_depth};
  double* o[4];
  for (int q = 0; q < 4; ++q) o[q] = w.out[q].get((std::size_t)W * nif);
  if (data.all_interfaces) {
    adrt_device_outputs out{data.flux_up ? data.flux_up : o[0],
                            data.flux_down ? data.flux_down : o[1],
                            data.mean_intensity ? data.mean_intensity : o[2],
                            data.flux_direct ? data.flux_direct : o[3], data.flux_divergence};
    check(adrt_solve_batch_from_coefficients(&c, &k, config.surface_temperature,
                                             config.top_temperature, &out, config.precision, ws,
                                             stream));
    return;
  }
  adrt_device_outputs out{o[0], o[1], o[2], o[3], nullptr};
  check(adrt_solve_batch_from_coefficients(&c, &k, config.surface_temperature,
                                           config.top_temperature, &out, config.precision, ws,
                                           stream));
  auto column = [&](double* dst, const double* src, int level) {
    if (dst)
      hip_check(hipMemcpy2DAsync(dst, sizeof(double), src + level, nif * sizeof(double),
                                 sizeof(double), W, hipMemcpyDeviceToDevice, stream),
                "hipMemcpy2DAsync");
  };
  column(data.flux_up, o[0], 0);
  column(data.flux_down, o[1], 0);
  column(data.flux_direct, o[3], L);
}

/// As above with a per-thread workspace (the reference's N <= 8 overload,
/// src_cuda/cuda_solver.cuh:156-162, allocates nothing per call either).
inline void solveBatchFromCoefficients(const BatchConfig& config, const RawDeviceData& data,
                                       hipStream_t stream = 0) {
  thread_local SolverWorkspaceGPU w;
  solveBatchFromCoefficients(config, data, w, stream);
}

struct HostResult {
  std::vector<double> flux_up;
  std::vector<double> flux_down;
  std::vector<double> flux_direct;
};

/// Host-side convenience (src_cuda/cuda_solver.cuh:91-98): copies the inputs
/// to the device, solves, copies the TOA results back, synchronises.  Host
/// vectors in the layouts of DeviceData (temperature and planck_levels
/// [nwav * nlev], or empty); a vector of another length throws
/// std::invalid_argument before anything is copied.
inline HostResult solveBatchHost(const BatchConfig& config, const std::vector<double>& delta_tau,
                                 const std::vector<double>& single_scat_albedo,
                                 const std::vector<double>& phase_moments,
                                 bool phase_moments_shared,
                                 const std::vector<double>& planck_levels,
                                 const std::vector<double>& temperature = {}) {
  using adrt::detail::DeviceBuffer;
  const std::size_t W = (std::size_t)std::max(config.num_wavenumbers, 0);
  const std::size_t L = (std::size_t)std::max(config.num_layers, 0);
  const std::size_t M = (std::size_t)std::max(config.num_moments_max, 0);
  // the device buffers are read with these extents: a shorter host vector
  // would be read past its end
  auto need = [](const std::vector<double>& v, std::size_t n, bool optional, const char* what) {
    if (optional && v.empty()) return;
    if (v.size() != n)
      throw std::invalid_argument(std::string("solveBatchHost: ") + what + " has " +
                                  std::to_string(v.size()) + " values, expected " +
                                  std::to_string(n));
  };
  need(delta_tau, W * L, false, "delta_tau");
  need(single_scat_albedo, W * L, false, "single_scat_albedo");
  need(phase_moments, (phase_moments_shared ? 1 : W) * L * M, false, "phase_moments");
  need(planck_levels, W * (L + 1), true, "planck_levels");
  need(temperature, W * (L + 1), true, "temperature");
  DeviceBuffer b[5], o[3];
  hipStream_t s = 0;
  DeviceData d;
  auto up = [&](DeviceBuffer& buf, const std::vector<double>& v) -> const double* {
    if (v.empty()) return nullptr;
    double* p = buf.get(v.size());
    adrt::detail::upload(p, v, s);
    return p;
  };
  d.delta_tau = up(b[0], delta_tau);
  d.single_scat_albedo = up(b[1], single_scat_albedo);
  d.phase_moments = up(b[2], phase_moments);
  d.phase_moments_shared = phase_moments_shared;
  d.planck_levels = up(b[3], planck_levels);
  d.temperature = up(b[4], temperature);
  d.flux_up = o[0].get(W);
  d.flux_down = o[1].get(W);
  d.flux_direct = o[2].get(W);
  solveBatch(config, d, s);
  HostResult r;
  adrt::detail::download(r.flux_up, d.flux_up, W, s);
  adrt::detail::download(r.flux_down, d.flux_down, W, s);
  adrt::detail::download(r.flux_direct, d.flux_direct, W, s);
  adrt::detail::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return r;
}

}  // namespace hip
}  // namespace adrt

#endif  // ADRT_HIP_HPP
