// Test driver of the C++ host layer (include/adrt_hip.hpp): reads one case
// from stdin, solves it on the GPU through adrt::solve, adrt::hip::solveBatch
// or adrt::hip::solveBatchFromCoefficients, prints the outputs as
// "name v0 v1 ..." lines, or "error <message>" when the configuration is
// rejected.
//
//   column: L N thermal delta_m diffusion from_bottom jacobian
//           albedo Bs Ts Bt Tt F0 mu0 wn_lo wn_hi
//           tau[L] ssa[L] M chi[L][M] nT T[nT] nB B[nB]
//   batch:  W L N M shared thermal delta_m diffusion from_bottom all_interfaces
//           (host: the same input through adrt::hip::solveBatchHost, TOA results)
//           albedo Bs Ts Bt Tt F0 mu0 wn_lo wn_hi scaling
//           tau[W][L] ssa[W][L] chi[(W)][L][M] nT T[nT] nB B[nB]
//   raw:    W L N M shared delta_m diffusion all_interfaces cloud
//           albedo Ts Tt F0 mu0
//           abs[L+1][W] sca[L+1][W] z[L+1] T[L+1] nu[W] chi[(W)][L][M] (cloud[L][W])
#include "adrt_hip.hpp"

#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

namespace {

std::vector<double> read_n(std::size_t n) {
  std::vector<double> v(n);
  for (double& x : v) std::cin >> x;
  return v;
}

void print(const char* name, const std::vector<double>& v) {
  std::printf("%s", name);
  for (double x : v) std::printf(" %.17g", x);
  std::printf("\n");
}

void print_jac(const char* name, const std::vector<std::vector<double>>& J) {
  std::vector<double> flat;
  for (const auto& row : J) flat.insert(flat.end(), row.begin(), row.end());
  print(name, flat);
}

int column() {
  adrt::ADConfig c;
  int thermal, dm, diff, fb, jac;
  std::cin >> c.num_layers >> c.num_quadrature >> thermal >> dm >> diff >> fb >> jac;
  c.use_thermal_emission = thermal;
  c.use_delta_m = dm;
  c.use_diffusion_lower_bc = diff;
  c.index_from_bottom = fb;
  c.compute_temperature_jacobian = jac;
  std::cin >> c.surface_albedo >> c.surface_emission >> c.surface_temperature >>
      c.top_emission >> c.top_temperature >> c.solar_flux >> c.solar_mu >> c.wavenumber_low >>
      c.wavenumber_high;
  const int L = c.num_layers;
  c.delta_tau = read_n(L);
  c.single_scat_albedo = read_n(L);
  int M;
  std::cin >> M;
  c.phase_function_moments.assign(L, {});
  for (int l = 0; l < L; ++l) c.phase_function_moments[l] = read_n(M);
  std::size_t nT, nB;
  std::cin >> nT;
  c.temperature = read_n(nT);
  std::cin >> nB;
  c.planck_levels = read_n(nB);
  adrt::RTOutput r;
  try {
    r = adrt::solve(c);
  } catch (const std::invalid_argument& e) {
    std::printf("error %s\n", e.what());
    return 0;
  }
  print("flux_up", r.flux_up);
  print("flux_down", r.flux_down);
  print("mean_intensity", r.mean_intensity);
  print("flux_divergence", r.flux_divergence);
  print("flux_direct", r.flux_direct);
  if (jac) {
    print_jac("flux_up_temperature_jac", r.flux_up_temperature_jac);
    print_jac("flux_down_temperature_jac", r.flux_down_temperature_jac);
    print_jac("mean_intensity_temperature_jac", r.mean_intensity_temperature_jac);
    print_jac("flux_divergence_temperature_jac", r.flux_divergence_temperature_jac);
  }
  return 0;
}

double* to_device(const std::vector<double>& v) {
  if (v.empty()) return nullptr;
  double* p = nullptr;
  adrt::detail::hip_check(hipMalloc((void**)&p, v.size() * sizeof(double)), "hipMalloc");
  adrt::detail::hip_check(
      hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  return p;
}

int batch(bool host) {
  adrt::hip::BatchConfig c;
  int shared, thermal, dm, diff, fb, all;
  std::cin >> c.num_wavenumbers >> c.num_layers >> c.num_quadrature >> c.num_moments_max >>
      shared >> thermal >> dm >> diff >> fb >> all;
  c.use_thermal_emission = thermal;
  c.use_delta_m = dm;
  c.use_diffusion_lower_bc = diff;
  c.index_from_bottom = fb;
  std::cin >> c.surface_albedo >> c.surface_emission >> c.surface_temperature >>
      c.top_emission >> c.top_temperature >> c.solar_flux >> c.solar_mu >> c.wavenumber_low >>
      c.wavenumber_high >> c.spectrum_scaling;
  const std::size_t W = c.num_wavenumbers, L = c.num_layers, M = c.num_moments_max;
  const auto tau = read_n(W * L), ssa = read_n(W * L);
  const auto mom = read_n((shared ? 1 : W) * L * M);
  std::size_t nT, nB;
  std::cin >> nT;
  const auto T = read_n(nT);
  std::cin >> nB;
  const auto B = read_n(nB);
  if (host) {  // adrt::hip::solveBatchHost: host vectors in, TOA results out
    adrt::hip::HostResult r;
    try {
      r = adrt::hip::solveBatchHost(c, tau, ssa, mom, shared, B, T);
    } catch (const std::invalid_argument& e) {
      std::printf("error %s\n", e.what());
      return 0;
    }
    print("flux_up", r.flux_up);
    print("flux_down", r.flux_down);
    print("flux_direct", r.flux_direct);
    return 0;
  }
  adrt::hip::DeviceData d;
  d.delta_tau = to_device(tau);
  d.single_scat_albedo = to_device(ssa);
  d.phase_moments = to_device(mom);
  d.phase_moments_shared = shared;
  d.temperature = to_device(T);
  d.planck_levels = to_device(B);
  d.all_interfaces = all;
  const std::size_t n = all ? W * (L + 1) : W;
  std::vector<double> zero(n, 0.0);
  d.flux_up = to_device(zero);
  d.flux_down = to_device(zero);
  d.flux_direct = to_device(zero);
  if (all) {
    d.mean_intensity = to_device(zero);
    d.flux_divergence = to_device(zero);
  }
  try {
    adrt::hip::solveBatch(c, d, 0);
  } catch (const std::invalid_argument& e) {
    std::printf("error %s\n", e.what());
    return 0;
  }
  adrt::detail::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  auto get = [&](const double* p) {
    std::vector<double> v(n);
    adrt::detail::hip_check(hipMemcpy(v.data(), p, n * sizeof(double), hipMemcpyDeviceToHost),
                            "hipMemcpy");
    return v;
  };
  print("flux_up", get(d.flux_up));
  print("flux_down", get(d.flux_down));
  print("flux_direct", get(d.flux_direct));
  if (all) {
    print("mean_intensity", get(d.mean_intensity));
    print("flux_divergence", get(d.flux_divergence));
  }
  return 0;
}

int raw() {
  adrt::hip::BatchConfig c;
  int shared, dm, diff, all, cloud;
  std::cin >> c.num_wavenumbers >> c.num_layers >> c.num_quadrature >> c.num_moments_max >>
      shared >> dm >> diff >> all >> cloud;
  c.use_delta_m = dm;
  c.use_diffusion_lower_bc = diff;
  std::cin >> c.surface_albedo >> c.surface_temperature >> c.top_temperature >> c.solar_flux >>
      c.solar_mu;
  const std::size_t W = c.num_wavenumbers, L = c.num_layers, M = c.num_moments_max;
  const auto a = read_n((L + 1) * W), s = read_n((L + 1) * W);
  const auto z = read_n(L + 1), T = read_n(L + 1), nu = read_n(W);
  const auto mom = read_n((shared ? 1 : W) * L * M);
  const auto cl = cloud ? read_n(L * W) : std::vector<double>();
  adrt::hip::RawDeviceData d;
  d.absorption_coeff = to_device(a);
  d.scattering_coeff = to_device(s);
  d.altitude = to_device(z);
  d.temperature = to_device(T);
  d.wavenumber = to_device(nu);
  d.phase_moments = to_device(mom);
  d.phase_moments_shared = shared;
  d.cloud_optical_depth = to_device(cl);
  d.all_interfaces = all;
  const std::size_t n = all ? W * (L + 1) : W;
  std::vector<double> zero(n, 0.0);
  d.flux_up = to_device(zero);
  d.flux_down = to_device(zero);
  d.flux_direct = to_device(zero);
  if (all) {
    d.mean_intensity = to_device(zero);
    d.flux_divergence = to_device(zero);
  }
  try {
    adrt::hip::solveBatchFromCoefficients(c, d, 0);
  } catch (const std::invalid_argument& e) {
    std::printf("error %s\n", e.what());
    return 0;
  }
  adrt::detail::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  auto get = [&](const double* p) {
    std::vector<double> v(n);
    adrt::detail::hip_check(hipMemcpy(v.data(), p, n * sizeof(double), hipMemcpyDeviceToHost),
                            "hipMemcpy");
    return v;
  };
  print("flux_up", get(d.flux_up));
  print("flux_down", get(d.flux_down));
  print("flux_direct", get(d.flux_direct));
  if (all) {
    print("mean_intensity", get(d.mean_intensity));
    print("flux_divergence", get(d.flux_divergence));
  }
  return 0;
}

}  // namespace

int main() {
  std::string mode;
  std::cin >> mode;
  try {
    if (mode == "column") return column();
    if (mode == "batch") return batch(false);
    if (mode == "host") return batch(true);
    if (mode == "raw") return raw();
  } catch (const std::exception& e) {
    std::printf("failure %s\n", e.what());
    return 2;
  }
  std::printf("failure unknown mode\n");
  return 2;
}
