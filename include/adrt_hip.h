/*
 * adrt_hip.h -- C-ABI of the MI355X-native batched adding-doubling solver.
 *
 * One call solves many independent columns (one per wavenumber) on one GPU.
 * All arrays are caller-owned device buffers in wavenumber-major layout,
 * array[w * stride + layer_or_level], TOA first unless index_from_bottom.
 * Every entry point returns an int status (0 = ADRT_OK) and never throws.
 * Work is enqueued on the given stream; nothing synchronises unless stated.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   adrt_solve_batch        <- adrt::cuda::solveBatch        src_cuda/cuda_solver.cuh:77-80
 *                              (FP64, every interface, not only TOA) and the
 *                              per-column adrt::solve        src/adding_doubling.hpp:258-269
 *   adrt_solve_batch_from_coefficients, adrt_device_coefficients
 *                           <- adrt::cuda::solveBatchFromCoefficients, RawDeviceData
 *                                                            src_cuda/cuda_solver.cuh:104-179
 *   adrt_batch_config       <- adrt::cuda::BatchConfig       src_cuda/cuda_solver.cuh:18-40
 *   adrt_device_inputs/outputs <- adrt::cuda::DeviceData     src_cuda/cuda_solver.cuh:45-68
 *   adrt_workspace_*        <- adrt::cuda::SolverWorkspaceGPU src_cuda/cuda_solver.cuh:140-156
 *                              and adrt::SolverWorkspace     src/workspace.hpp:23-48
 *   adrt_gauss_legendre     <- adrt::gaussLegendre           src/adding_doubling.hpp:277-280
 *   adrt_planck / _deriv    <- adrt::planckFunction(Deriv)   src/planck.hpp, src/planck.cpp:35-177
 *   status codes 1..17      <- the std::invalid_argument cases of
 *                              adrt::ADConfig::validate     src/adding_doubling.hpp:123-174
 *
 *   adrt_source_jacobian_batch <- adrt::computeTemperatureJacobian (the Planck-value part)
 *                                                            src/solver.cpp:117-341
 *   adrt_planck_levels      <- the device Planck             src_cuda/cuda_planck.cuh:33-122
 *   adrt_validate_inputs    <- the tau / omega range checks of adrt::ADConfig::validate,
 *                              on the device
 *
 * Without a counterpart (the reference has no derivatives): adrt_adjoint_batch
 * and its workspace flag, adrt_raw_optics / adrt_raw_optics_adjoint (the
 * raw-coefficient optics and their reverse), adrt_profile / adrt_profile_read;
 * adrt_solve_batch_noalloc (for callers that cannot allocate).
 */
#ifndef ADRT_HIP_H
#define ADRT_HIP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ADRT_OK 0
#define ADRT_ERR_NUM_LAYERS 1          /* num_layers must be > 0 */
#define ADRT_ERR_NUM_QUADRATURE 2      /* num_quadrature must be >= 2 */
#define ADRT_ERR_DELTA_TAU 3           /* delta_tau missing */
#define ADRT_ERR_SSA 4                 /* single_scat_albedo missing */
#define ADRT_ERR_MOMENTS 5             /* phase moments missing / num_moments < 1 */
#define ADRT_ERR_NEGATIVE_TAU 6        /* delta_tau[l] < 0 (checked by adrt_validate_inputs) */
#define ADRT_ERR_SSA_RANGE 7           /* single_scat_albedo out of [0,1] (idem) */
#define ADRT_ERR_SURFACE_ALBEDO 8      /* surface_albedo out of [0,1] */
#define ADRT_ERR_SOLAR_MU 9            /* solar_mu must be in (0,1] when solar_flux > 0 */
#define ADRT_ERR_TEMPERATURE_SIZE 10
#define ADRT_ERR_TEMPERATURE 11
#define ADRT_ERR_WAVENUMBER_RANGE 12
#define ADRT_ERR_SURFACE_TEMPERATURE 13
#define ADRT_ERR_TOP_TEMPERATURE 14
#define ADRT_ERR_DIFFUSION_BC 15       /* diffusion lower BC needs Planck levels */
#define ADRT_ERR_PLANCK_ARGS 16
#define ADRT_ERR_QUADRATURE_MAX 17     /* num_quadrature > ADRT_MAX_QUADRATURE */
#define ADRT_ERR_NUM_WAVENUMBERS 18    /* num_wavenumbers must be >= 0 */
#define ADRT_ERR_OUTPUTS 19            /* a required output pointer is NULL */
#define ADRT_ERR_WORKSPACE 20          /* workspace missing or made for another shape */
#define ADRT_ERR_HIP 21                /* a HIP runtime call failed */
#define ADRT_ERR_PRECISION 22          /* unknown precision value */
#define ADRT_ERR_NULL 23               /* NULL argument */
#define ADRT_ERR_ADJOINT 24            /* workspace not created with ADRT_WORKSPACE_ADJOINT */
#define ADRT_ERR_DOUBLING_STEPS 25     /* a layer needs more doubling steps than the reverse pass holds */
#define ADRT_ERR_SCRATCH 26            /* adrt_solve_batch_noalloc: one column exceeds its scratch */

#define ADRT_MAX_QUADRATURE 32

typedef enum adrt_precision {
  ADRT_PRECISION_F64 = 0,   /* plain form: FP64 throughout */
  ADRT_PRECISION_MIXED = 1, /* FP32 doubling (FP64 start, FP64-refined solves), FP64 adding and
                               fluxes; layers with an unstable doubling start and their
                               neighbours in FP64; reverse doubling in FP32 likewise.  Padded
                               orders 16 (9..16, or any order on a reverse-pass workspace) and
                               32 (17..32, forward only) */
  ADRT_PRECISION_MIXED_REVERSE = 2 /* the FP64 forward (outputs identical to ADRT_PRECISION_F64)
                                      with the reverse pass's doubling recomputation and its
                                      reverse in FP32, routed as for ADRT_PRECISION_MIXED */
} adrt_precision;

/* Scalars shared by every column (names as BatchConfig). */
typedef struct adrt_batch_config {
  int num_wavenumbers;
  int num_layers;
  int num_quadrature;          /* 2..32; 2,4,8,16,32 native, others padded */
  int num_moments;             /* moments per layer (M) */
  int use_delta_m;             /* delta-M when M > 2N (ADConfig::use_delta_m) */
  int use_diffusion_lower_bc;  /* diffusion lower boundary (needs planck_levels) */
  int index_from_bottom;       /* inputs/outputs BOA first */
  double surface_albedo;       /* Lambertian albedo in [0,1] */
  double surface_emission;     /* used when inputs.surface_emission == NULL */
  double top_emission;         /* used when inputs.top_emission == NULL */
  double solar_flux;           /* collimated flux at TOA (0 = no beam) */
  double solar_mu;             /* cosine of the solar zenith angle, (0,1] */
  double spectrum_scaling;     /* factor on flux_up (1.0 = none) */
} adrt_batch_config;

/* Caller-owned device inputs (FP64). */
typedef struct adrt_device_inputs {
  const double* delta_tau;           /* [W][L] */
  const double* single_scat_albedo;  /* [W][L] */
  const double* phase_moments;       /* [L][M] if shared, else [W][L][M] */
  int phase_moments_shared;
  const double* planck_levels;       /* [W][L+1] Planck at levels, or NULL (no thermal) */
  const double* surface_emission;    /* [W] or NULL */
  const double* top_emission;        /* [W] or NULL */
} adrt_device_inputs;

/* Caller-owned device outputs (FP64), [W][L+1] at every interface. */
typedef struct adrt_device_outputs {
  double* flux_up;
  double* flux_down;
  double* mean_intensity;   /* includes the direct-beam term F_dir/(4 pi mu0) */
  double* flux_direct;
  double* flux_divergence;  /* may be NULL */
} adrt_device_outputs;

/* Cotangents of the outputs (dLoss/d output), [W][L+1] each in the layout of
 * adrt_device_outputs; NULL = zero. */
typedef struct adrt_device_cotangents {
  const double* flux_up;
  const double* flux_down;
  const double* mean_intensity;
  const double* flux_direct;
  const double* flux_divergence;
} adrt_device_cotangents;

/* Gradients written by adrt_adjoint_batch, in the layout of the inputs; NULL =
 * not wanted.  Every non-NULL array is overwritten. */
typedef struct adrt_device_gradients {
  double* delta_tau;           /* [W][L] */
  double* single_scat_albedo;  /* [W][L] */
  double* phase_moments;       /* [L][M] if the moments are shared (summed over W), else [W][L][M] */
  double* planck_levels;       /* [W][L+1] */
  double* surface_emission;    /* [W] (per column, whether or not the input was per column) */
  double* top_emission;        /* [W] */
  double* surface_albedo;      /* [W] per-column parts; the albedo gradient is their sum */
} adrt_device_gradients;

typedef struct adrt_workspace adrt_workspace;

/* Workspace flags. */
#define ADRT_WORKSPACE_ADJOINT 1  /* keep what the reverse pass needs */
#define ADRT_WORKSPACE_JACOBIAN 2 /* room for adrt_source_jacobian_batch (not with ADJOINT) */

/* Create a workspace for columns of this shape (num_layers, num_quadrature,
 * num_moments).  max_bytes bounds its device memory (0 = 4 GiB); the batch is
 * processed in chunks of wavenumbers that fit.  Allocates device memory. */
int adrt_workspace_create(const adrt_batch_config* cfg, size_t max_bytes,
                          adrt_workspace** ws);
/* As adrt_workspace_create, with flags (ADRT_WORKSPACE_ADJOINT).  With the
 * adjoint flag the columns are solved at quadrature order 16 or, for orders
 * 17..32, 32 (lower orders padded with zero-weight nodes, which leaves the
 * results unchanged) and the
 * workspace also holds interface intensities, per-layer gradient terms and
 * the doubling checkpoints of the reverse pass. */
int adrt_workspace_create_ex(const adrt_batch_config* cfg, size_t max_bytes, int flags,
                             adrt_workspace** ws);
int adrt_workspace_destroy(adrt_workspace* ws);
/* Device bytes held and wavenumbers per chunk. */
int adrt_workspace_info(const adrt_workspace* ws, size_t* bytes, int* chunk_wavenumbers);

/* Forward solve of num_wavenumbers columns.  Asynchronous on stream. */
int adrt_solve_batch(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                     const adrt_device_outputs* out, adrt_precision precision,
                     adrt_workspace* ws, hipStream_t stream);

/* Derivatives of the outputs with respect to the thermal sources, FP64:
 * array[(w * (L + 3) + u) * (L + 1) + i] = d output(w, interface i) / d source u,
 * u = 0..L the Planck value of level u, u = L + 1 the surface emission, u = L + 2
 * the top emission (levels and interfaces in the caller's order). */
typedef struct adrt_device_jacobians {
  double* flux_up;
  double* flux_down;
  double* mean_intensity;
  double* flux_divergence;  /* may be NULL */
} adrt_device_jacobians;

/* Source Jacobians of num_wavenumbers columns (FP64, any quadrature order):
 * replaces the forward-mode adrt::computeTemperatureJacobian
 * (src/solver.cpp:117-341), which differentiates the adding recursion and the
 * interface solves with respect to every level's Planck value, the surface
 * and the top emission, reusing the factorisations of the solve.  R and T do
 * not depend on the sources and every output is affine in them, so each
 * column's doubling runs once (no beam, no sources), its layers keep the
 * thermal source basis (s_up = y B_bar + z B_d, s_dn = y B_bar - z B_d), and the
 * adding sweeps and interface solves run for L + 3 pseudo-columns per column,
 * each with one unit source.  Temperature Jacobians follow by the chain rule
 * with dB/dT (adrt_planck_levels with derivative != 0).  The inputs'
 * planck_levels, surface_emission and top_emission are not read, nor is the
 * beam.  Needs a workspace created with ADRT_WORKSPACE_JACOBIAN (the batch runs
 * in chunks of the columns that fit it).  Asynchronous on stream. */
int adrt_source_jacobian_batch(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                               const adrt_device_jacobians* jac, adrt_workspace* ws,
                               hipStream_t stream);

/* Scratch held by adrt_solve_batch_noalloc on each GPU it runs on, for the
 * life of the process (a file-scope device array of the library). */
#define ADRT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* The FP64 forward solve of adrt_solve_batch without a workspace and without
 * allocating anything: no hipMalloc / hipFree, no allocating library call, no
 * synchronisation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference, whose solveBatch mallocs and frees its
 * temporaries inside every call (src_cuda/cuda_solver.cu:899-912).  Its scratch
 * -- quadrature tables, per-chunk optics, layer records, base composites --
 * is the library's ADRT_NOALLOC_SCRATCH_BYTES array on the current GPU; the
 * batch runs in chunks of wavenumbers that fit it, and a column shape of which
 * not even one column fits returns ADRT_ERR_SCRATCH (at quadrature order 16,
 * about 0.87 MB per column of 100 layers, so 76 columns per chunk).  Results
 * equal adrt_solve_batch's with precision ADRT_PRECISION_F64.  Every call on a
 * GPU shares that GPU's array and leaves its contents behind (nothing in it
 * needs to start at zero): calls that overlap on one GPU -- on two streams, or
 * from two threads -- race, so serialise them.  Asynchronous on stream. */
int adrt_solve_batch_noalloc(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                             const adrt_device_outputs* out, hipStream_t stream);

/* Token of the last adrt_solve_batch on this workspace (0 before the first),
 * unique among all forward solves of the process, on any workspace; pass the
 * value read right after a forward solve to adrt_adjoint_batch to let it
 * reuse that solve's layer records and interface intensities. */
int adrt_workspace_generation(const adrt_workspace* ws, unsigned long long* generation);

/* Reverse pass: gradients of sum(cotangent * output) with respect to the
 * inputs, for the columns solved by adrt_solve_batch with the same cfg and in.
 * out must hold that solve's outputs (mean_intensity is read; when the
 * forward is recomputed they are written again with the same values).  The
 * forward state kept in the workspace is reused when forward_generation equals
 * the workspace generation and the whole batch fit one chunk; otherwise each
 * chunk's forward is recomputed (checkpointing: layer records and interface
 * intensities are kept per chunk, doubling steps are always recomputed).
 * The reference has no adjoint; its derivative feature is the forward-mode
 * temperature Jacobian of src/solver.cpp:117-341.  Every quadrature order;
 * orders 17..32 in ADRT_PRECISION_F64 only (else ADRT_ERR_PRECISION).
 * Asynchronous on stream; adrt_adjoint_status reports layers that exceeded
 * the checkpoint depth (64 doubling steps up to order 16, 48 above). */
int adrt_adjoint_batch(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                       const adrt_device_outputs* out, const adrt_device_cotangents* cot,
                       const adrt_device_gradients* grad, adrt_precision precision,
                       unsigned long long forward_generation, adrt_workspace* ws,
                       hipStream_t stream);
/* Synchronises the stream; ADRT_ERR_DOUBLING_STEPS if a layer of a reverse
 * pass since the last call needed more doubling steps than the checkpoint slab
 * holds (its gradients were written as zero), else ADRT_OK.  Clears the flag. */
int adrt_adjoint_status(adrt_workspace* ws, hipStream_t stream);

/* Raw retrieval inputs (the BeAR format), caller-owned device memory, FP64.
 * Coefficient arrays are level-major with the bottom of the atmosphere first,
 * array[level * num_wavenumbers + w]; the solver takes layer l (TOA = 0) from
 * levels L-1-l and L-l.  Replaces adrt::cuda::RawDeviceData
 * (src_cuda/cuda_solver.cuh:104-134). */
typedef struct adrt_device_coefficients {
  const double* absorption_coeff;     /* [L+1][W] (cm^-1), BOA first */
  const double* scattering_coeff;     /* [L+1][W] (cm^-1), BOA first */
  const double* altitude;             /* [L+1] (cm), BOA first */
  const double* temperature;          /* [L+1] (K), BOA first */
  const double* wavenumber;           /* [W] (cm^-1) */
  const double* phase_moments;        /* [L][M] if shared, else [W][L][M]; layers TOA first */
  int phase_moments_shared;
  const double* cloud_optical_depth;  /* [L][W] layer-major, BOA first, or NULL */
} adrt_device_coefficients;

/* Solve from raw coefficients (adrt::cuda::solveBatchFromCoefficients,
 * src_cuda/cuda_solver.cuh:164-179, src_cuda/cuda_solver.cu:131-221,1131-1353):
 * per wavenumber and layer, tau = dz (ext_bot + ext_top) / 2 + cloud tau
 * (clamped at 0) with ext = absorption + scattering, omega = dz (scat_bot +
 * scat_top) / 2 / tau (0 where tau = 0); level Planck values are the
 * single-wavenumber B(T, nu) = c1 nu^3 / (exp(c2 nu / T) - 1) (0 below 1e-5 K);
 * the surface emits B(surface_temperature) if surface_temperature >= 0, else
 * B at the bottom level, and the top boundary receives B(top_temperature) if
 * top_temperature >= 0, else B at the top level.  cfg->surface_emission,
 * top_emission and index_from_bottom are not used: outputs are [W][L+1] at
 * every interface, TOA first.  Asynchronous on stream. */
int adrt_solve_batch_from_coefficients(const adrt_batch_config* cfg,
                                       const adrt_device_coefficients* coef,
                                       double surface_temperature, double top_temperature,
                                       const adrt_device_outputs* out, adrt_precision precision,
                                       adrt_workspace* ws, hipStream_t stream);

/* The optics the raw entry prepares, for all num_wavenumbers columns at once
 * into caller buffers: delta_tau and single_scat_albedo [W][L], planck_levels
 * [W][L+1], surface_emission and top_emission [W], TOA first; the rules of
 * adrt_solve_batch_from_coefficients.  Feeding them to adrt_solve_batch gives
 * that entry's result.  No counterpart of its own in the reference (its
 * preprocessRawKernel, src_cuda/cuda_solver.cu:1131-1182, writes the same
 * arrays into its workspace).  Asynchronous on stream. */
int adrt_raw_optics(int num_wavenumbers, int num_layers, const adrt_device_coefficients* coef,
                    double surface_temperature, double top_temperature, double* delta_tau,
                    double* single_scat_albedo, double* planck_levels, double* surface_emission,
                    double* top_emission, hipStream_t stream);

/* Cotangents of adrt_raw_optics' outputs (layouts as there); NULL = zero. */
typedef struct adrt_raw_optics_cotangents {
  const double* delta_tau;
  const double* single_scat_albedo;
  const double* planck_levels;
  const double* surface_emission;
  const double* top_emission;
} adrt_raw_optics_cotangents;

/* Gradients written by adrt_raw_optics_adjoint, in the layouts of
 * adrt_device_coefficients; NULL = not wanted.  temperature [L+1] and the two
 * boundary temperatures (one double each, 0 when the temperature is not set)
 * are sums over the wavenumbers. */
typedef struct adrt_raw_gradients {
  double* absorption_coeff;      /* [L+1][W] */
  double* scattering_coeff;      /* [L+1][W] */
  double* cloud_optical_depth;   /* [L][W] (written only if the coefficients have clouds) */
  double* temperature;           /* [L+1] */
  double* surface_temperature;   /* [1] */
  double* top_temperature;       /* [1] */
} adrt_raw_gradients;

/* Reverse of adrt_raw_optics: with the cotangents of the optics (from
 * adrt_adjoint_batch run on them), the gradients w.r.t. the raw coefficients
 * and temperatures.  Clamped layers (tau <= 0) pass nothing back; altitudes and
 * wavenumbers get no gradient.  No counterpart in the reference (it has no
 * derivatives of the raw path).  Asynchronous on stream; deterministic. */
int adrt_raw_optics_adjoint(int num_wavenumbers, int num_layers,
                            const adrt_device_coefficients* coef, double surface_temperature,
                            double top_temperature, const adrt_raw_optics_cotangents* cot,
                            const adrt_raw_gradients* grad, hipStream_t stream);

/* Checks delta_tau >= 0 and single_scat_albedo in [0,1] on the device and
 * returns ADRT_ERR_NEGATIVE_TAU / ADRT_ERR_SSA_RANGE (synchronises the stream). */
int adrt_validate_inputs(const adrt_batch_config* cfg, const adrt_device_inputs* in,
                         adrt_workspace* ws, hipStream_t stream);

/* Band Planck function (or dB/dT when derivative != 0) at every level of every
 * column, on the device: out[w][l] = B(wn_lo[w], wn_hi[w], T), T = temperature[l]
 * if temperature_shared else temperature[w][l].  Synchronises the stream;
 * returns ADRT_ERR_PLANCK_ARGS if any (band, T) is invalid.
 * Replaces the per-level planckFunction calls of the column solver
 * (src/solver.cpp:385-419) and the device Planck of src_cuda/cuda_planck.cuh:33-122. */
int adrt_planck_levels(const double* wn_lo, const double* wn_hi, const double* temperature,
                       int temperature_shared, int num_wavenumbers, int num_levels,
                       int derivative, double* out, adrt_workspace* ws, hipStream_t stream);

/* Opt-in kernel timing (no counterpart in the reference, whose benchmark times
 * whole calls, tests/benchmark_cuda.cu:209-222).  While enabled, every launch
 * of adrt_solve_batch on this workspace is bracketed by HIP events on the
 * caller's stream.  adrt_profile_read synchronises those events and returns
 * the summed milliseconds and launch counts per kernel kind since the last
 * read (ADRT_PROFILE_KINDS entries): [0] prep, [1] layer setup + doubling,
 * [2] adding + interfaces, [3] reverse adding + interfaces, [4] reverse layer
 * doubling (recomputed doubling + its reverse), [5] reverse column terms +
 * reductions, [6] reverse of the doubling start (phase matrices, Legendre sums).
 * Enabling allocates HIP events (host-side objects) on first use. */
#define ADRT_PROFILE_KINDS 7
int adrt_profile(adrt_workspace* ws, int enable);
int adrt_profile_read(adrt_workspace* ws, double* ms, int* launches);

/* Host utilities. */
int adrt_gauss_legendre(int n, double* nodes, double* weights);
int adrt_planck(double wnumlo, double wnumhi, double temp, double* out);
int adrt_planck_deriv(double wnumlo, double wnumhi, double temp, double* out);
const char* adrt_status_string(int status);
/* Library version string. */
const char* adrt_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ADRT_HIP_H */
