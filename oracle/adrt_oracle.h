/*
 * adrt_oracle.h -- FP64 CPU restatement of the AddingDoublingRT column solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product (addingdoublingrt_amd/) never links,
 * calls or falls back to it.
 *
 * Every routine restates the reference CPU solver (src/ of the reference) in
 * plain C with a runtime quadrature order; the file:line each function follows
 * is given at its definition in adrt_oracle.c.
 */
#ifndef ADRT_ORACLE_H
#define ADRT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_N 64

/* Column configuration: mirrors adrt::ADConfig (src/adding_doubling.hpp:63-94).
 * phase_moments is row-major [num_layers][num_moments]. */
typedef struct orc_config {
  int num_layers;
  int num_quadrature;
  int num_moments;
  int use_thermal_emission;
  int use_delta_m;
  int use_diffusion_lower_bc;
  int index_from_bottom;
  int compute_temperature_jacobian;
  double surface_albedo;
  double surface_emission;
  double surface_temperature;
  double top_emission;
  double top_temperature;
  double solar_flux;
  double solar_mu;
  double wavenumber_low;
  double wavenumber_high;
  const double* delta_tau;          /* [L] */
  const double* single_scat_albedo; /* [L] */
  const double* temperature;        /* [L+1] or NULL */
  const double* planck_levels;      /* [L+1] or NULL */
  const double* phase_moments;      /* [L][M] */
} orc_config;

/* Outputs: mirrors adrt::RTOutput (src/adding_doubling.hpp:32-50).
 * Interface arrays are [L+1]; Jacobians are [(L+1)][(L+2)] row-major.
 * Any pointer may be NULL. */
typedef struct orc_output {
  double* flux_up;
  double* flux_down;
  double* mean_intensity;
  double* flux_divergence;
  double* flux_direct;
  double* jac_flux_up;
  double* jac_flux_down;
  double* jac_mean_intensity;
  double* jac_flux_divergence;
} orc_output;

/* Batched column set in the product's wavenumber-major layout (planck_levels
 * mode).  Mirrors the batched entry of the HIP library so that both sides are
 * fed the same buffers. */
typedef struct orc_batch {
  int num_wavenumbers;
  int num_layers;
  int num_quadrature;
  int num_moments;
  int phase_moments_shared;   /* 1: moments [L][M]; 0: [W][L][M] */
  int use_delta_m;
  int use_diffusion_lower_bc;
  int index_from_bottom;
  double surface_albedo;
  double surface_emission;    /* used when surface_emission_w == NULL */
  double top_emission;        /* used when top_emission_w == NULL */
  double solar_flux;
  double solar_mu;
  double spectrum_scaling;    /* multiplies flux_up (1.0 = none) */
  const double* delta_tau;          /* [W][L] */
  const double* single_scat_albedo; /* [W][L] */
  const double* phase_moments;      /* [L][M] or [W][L][M] */
  const double* planck_levels;      /* [W][L+1] or NULL (no thermal source) */
  const double* surface_emission_w; /* [W] or NULL */
  const double* top_emission_w;     /* [W] or NULL */
  double* flux_up;                  /* [W][L+1] */
  double* flux_down;
  double* mean_intensity;
  double* flux_direct;
  double* flux_divergence;          /* may be NULL */
} orc_batch;

/* Returns 0 on success, a positive error code (matching the validation
 * message written to err) on invalid input. */
int orc_solve(const orc_config* cfg, orc_output* out, char* err, int errlen);
int orc_solve_batch(const orc_batch* b, int num_threads, char* err, int errlen);

/* Building blocks (exported for unit tests). */
void orc_gauss_legendre(int n, double* nodes, double* weights);
void orc_legendre_table(int L, const double* x, int nx, double* Pl /* [L][nx] */);
void orc_phase_matrices(const double* chi, int L, const double* mu,
                        const double* wt, int n, double* Ppp, double* Ppm);
void orc_solar_phase_vectors(const double* chi, int L, const double* mu,
                             const double* wt, int n, double mu0,
                             double* p_plus, double* p_minus);
int orc_planck(double wnumlo, double wnumhi, double temp, double* out);
int orc_planck_deriv(double wnumlo, double wnumhi, double temp, double* out);
int orc_doubling_count(double tau, double omega);

/* One homogeneous layer through the reference doubling routine.
 * R, T: [n][n]; vectors [n].  Returns is_scattering. */
int orc_doubling(int n, double tau, double omega, double B_top, double B_bottom,
                 const double* Ppp, const double* Ppm, const double* mu,
                 const double* wt, double solar_flux, double solar_mu,
                 double tau_cumulative, const double* p_plus_solar,
                 const double* p_minus_solar, double* R, double* T,
                 double* s_up, double* s_down, double* s_up_solar,
                 double* s_down_solar, double* j_p, double* j_q);

#ifdef __cplusplus
}
#endif

#endif
