"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference CPU column solver
(``src/solver.cpp``, ``src/doubling.hpp``, ``src/adding.hpp``,
``src/phase_matrix.hpp``, ``src/quadrature.cpp``, ``src/planck.cpp`` of the
reference).  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may use it, and only as the checker: the product library
never calls it.

``solve(**cfg)`` mirrors ``adrt::solve(const ADConfig&)`` and
``solve_batch(...)`` feeds the same wavenumber-major buffers as the HIP entry.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_adrt.so")
_LIB_LD_PATH = os.path.join(_HERE, "liboracle_adrt_ld.so")
_lib = None
_lib_ld = None

_dp = C.POINTER(C.c_double)


class _Config(C.Structure):
    _fields_ = [
        ("num_layers", C.c_int), ("num_quadrature", C.c_int), ("num_moments", C.c_int),
        ("use_thermal_emission", C.c_int), ("use_delta_m", C.c_int),
        ("use_diffusion_lower_bc", C.c_int), ("index_from_bottom", C.c_int),
        ("compute_temperature_jacobian", C.c_int),
        ("surface_albedo", C.c_double), ("surface_emission", C.c_double),
        ("surface_temperature", C.c_double), ("top_emission", C.c_double),
        ("top_temperature", C.c_double), ("solar_flux", C.c_double),
        ("solar_mu", C.c_double), ("wavenumber_low", C.c_double),
        ("wavenumber_high", C.c_double),
        ("delta_tau", _dp), ("single_scat_albedo", _dp), ("temperature", _dp),
        ("planck_levels", _dp), ("phase_moments", _dp),
    ]


class _Output(C.Structure):
    _fields_ = [(n, _dp) for n in (
        "flux_up", "flux_down", "mean_intensity", "flux_divergence", "flux_direct",
        "jac_flux_up", "jac_flux_down", "jac_mean_intensity", "jac_flux_divergence")]


class _Batch(C.Structure):
    _fields_ = [
        ("num_wavenumbers", C.c_int), ("num_layers", C.c_int), ("num_quadrature", C.c_int),
        ("num_moments", C.c_int), ("phase_moments_shared", C.c_int),
        ("use_delta_m", C.c_int), ("use_diffusion_lower_bc", C.c_int),
        ("index_from_bottom", C.c_int),
        ("surface_albedo", C.c_double), ("surface_emission", C.c_double),
        ("top_emission", C.c_double), ("solar_flux", C.c_double), ("solar_mu", C.c_double),
        ("spectrum_scaling", C.c_double),
        ("delta_tau", _dp), ("single_scat_albedo", _dp), ("phase_moments", _dp),
        ("planck_levels", _dp), ("surface_emission_w", _dp), ("top_emission_w", _dp),
        ("flux_up", _dp), ("flux_down", _dp), ("mean_intensity", _dp),
        ("flux_direct", _dp), ("flux_divergence", _dp),
    ]


class OracleError(ValueError):
    """Raised for the validation failures the reference reports as std::invalid_argument."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def build() -> str:
    """Compile the oracle with its Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib(extended: bool = False):
    """The FP64 oracle, or (extended=True) the same code in x87 long double."""
    global _lib, _lib_ld
    if extended:
        if _lib_ld is None:
            if not os.path.exists(_LIB_LD_PATH):
                build()
            L = C.CDLL(_LIB_LD_PATH)
            L.orc_solve.argtypes = [C.POINTER(_Config), C.POINTER(_Output), C.c_char_p, C.c_int]
            L.orc_solve_batch.argtypes = [C.POINTER(_Batch), C.c_int, C.c_char_p, C.c_int]
            _lib_ld = L
        return _lib_ld
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_solve.argtypes = [C.POINTER(_Config), C.POINTER(_Output), C.c_char_p, C.c_int]
        L.orc_solve_batch.argtypes = [C.POINTER(_Batch), C.c_int, C.c_char_p, C.c_int]
        L.orc_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
        L.orc_legendre_table.argtypes = [C.c_int, _dp, C.c_int, _dp]
        L.orc_phase_matrices.argtypes = [_dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
        L.orc_solar_phase_vectors.argtypes = [_dp, C.c_int, _dp, _dp, C.c_int, C.c_double,
                                              _dp, _dp]
        L.orc_planck.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_planck_deriv.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_doubling_count.argtypes = [C.c_double, C.c_double]
        L.orc_doubling.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                   _dp, _dp] + [_dp] * 8
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _arr(x, n=None):
    if x is None:
        return None
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if n is not None and a.size != n:
        raise ValueError("array size mismatch")
    return a


def gauss_legendre(n: int):
    mu = np.zeros(n)
    wt = np.zeros(n)
    lib().orc_gauss_legendre(n, _ptr(mu), _ptr(wt))
    return mu, wt


def legendre_table(L: int, x):
    x = _arr(x)
    out = np.zeros((L, x.size))
    lib().orc_legendre_table(L, _ptr(x), x.size, _ptr(out))
    return out


def phase_matrices(chi, mu, wt):
    chi, mu, wt = _arr(chi), _arr(mu), _arr(wt)
    n = mu.size
    Ppp = np.zeros((n, n))
    Ppm = np.zeros((n, n))
    lib().orc_phase_matrices(_ptr(chi), chi.size, _ptr(mu), _ptr(wt), n, _ptr(Ppp), _ptr(Ppm))
    return Ppp, Ppm


def solar_phase_vectors(chi, mu, wt, mu0):
    chi, mu, wt = _arr(chi), _arr(mu), _arr(wt)
    n = mu.size
    pp = np.zeros(n)
    pm = np.zeros(n)
    lib().orc_solar_phase_vectors(_ptr(chi), chi.size, _ptr(mu), _ptr(wt), n, mu0,
                                  _ptr(pp), _ptr(pm))
    return pp, pm


def planck(lo: float, hi: float, T: float) -> float:
    out = C.c_double()
    rc = lib().orc_planck(lo, hi, T, C.byref(out))
    if rc:
        raise OracleError(rc, "planckFunction: invalid arguments")
    return out.value


def planck_deriv(lo: float, hi: float, T: float) -> float:
    out = C.c_double()
    rc = lib().orc_planck_deriv(lo, hi, T, C.byref(out))
    if rc:
        raise OracleError(rc, "planckFunctionDeriv: invalid arguments")
    return out.value


def doubling_count(tau: float, omega: float) -> int:
    return lib().orc_doubling_count(tau, omega)


def doubling(n, tau, omega, B_top, B_bottom, Ppp, Ppm, mu, wt, solar_flux=0.0,
             solar_mu=0.0, tau_cumulative=0.0, p_plus=None, p_minus=None):
    """One homogeneous layer (``doubling<N>``, src/doubling.hpp:29-230)."""
    Ppp, Ppm, mu, wt = _arr(Ppp), _arr(Ppm), _arr(mu), _arr(wt)
    p_plus, p_minus = _arr(p_plus), _arr(p_minus)
    R = np.zeros((n, n))
    T = np.zeros((n, n))
    vec = [np.zeros(n) for _ in range(6)]
    scat = lib().orc_doubling(n, tau, omega, B_top, B_bottom, _ptr(Ppp), _ptr(Ppm), _ptr(mu),
                              _ptr(wt), solar_flux, solar_mu, tau_cumulative, _ptr(p_plus),
                              _ptr(p_minus), _ptr(R), _ptr(T), *[_ptr(v) for v in vec])
    names = ("s_up", "s_down", "s_up_solar", "s_down_solar", "j_p", "j_q")
    out = dict(zip(names, vec))
    out.update(R=R, T=T, is_scattering=bool(scat))
    return out


def solve(num_layers, num_quadrature=8, delta_tau=None, single_scat_albedo=None,
          phase_function_moments=None, temperature=None, planck_levels=None,
          use_thermal_emission=False, use_delta_m=False, use_diffusion_lower_bc=False,
          index_from_bottom=False, compute_temperature_jacobian=False,
          surface_albedo=0.0, surface_emission=0.0, surface_temperature=-1.0,
          top_emission=0.0, top_temperature=-1.0, solar_flux=0.0, solar_mu=1.0,
          wavenumber_low=0.0, wavenumber_high=0.0, extended=False):
    """One column, mirroring ``adrt::solve`` (src/solver.cpp:348-722).

    Returns a dict of numpy arrays with the ``RTOutput`` field names.
    Raises OracleError for the cases the reference rejects in ``validate``.
    """
    L = int(num_layers)
    tau = _arr(delta_tau)
    ssa = _arr(single_scat_albedo)
    mom = None
    M = 0
    if phase_function_moments is not None:
        mom = np.ascontiguousarray(np.atleast_2d(np.asarray(phase_function_moments, np.float64)))
        M = mom.shape[1]
    temp = _arr(temperature)
    plev = _arr(planck_levels)
    cfg = _Config(
        num_layers=L, num_quadrature=int(num_quadrature), num_moments=M,
        use_thermal_emission=int(bool(use_thermal_emission)), use_delta_m=int(bool(use_delta_m)),
        use_diffusion_lower_bc=int(bool(use_diffusion_lower_bc)),
        index_from_bottom=int(bool(index_from_bottom)),
        compute_temperature_jacobian=int(bool(compute_temperature_jacobian)),
        surface_albedo=surface_albedo, surface_emission=surface_emission,
        surface_temperature=surface_temperature, top_emission=top_emission,
        top_temperature=top_temperature, solar_flux=solar_flux, solar_mu=solar_mu,
        wavenumber_low=wavenumber_low, wavenumber_high=wavenumber_high,
        delta_tau=_ptr(tau), single_scat_albedo=_ptr(ssa), temperature=_ptr(temp),
        planck_levels=_ptr(plev), phase_moments=_ptr(mom))
    # The reference checks array sizes (adding_doubling.hpp:128-133,150-151).
    if tau is not None and tau.size != L:
        cfg.delta_tau = None
    if ssa is not None and ssa.size != L:
        cfg.single_scat_albedo = None
    if mom is not None and mom.shape[0] != L:
        cfg.phase_moments = None
    if temp is not None and temp.size != L + 1:
        cfg.temperature = None
    if plev is not None and plev.size != L + 1:
        cfg.planck_levels = None
    nif = max(L, 0) + 1
    res = {k: np.zeros(nif) for k in ("flux_up", "flux_down", "mean_intensity",
                                        "flux_divergence", "flux_direct")}
    jac = {}
    out = _Output(**{k: _ptr(v) for k, v in res.items()})
    if compute_temperature_jacobian:
        for k in ("flux_up", "flux_down", "mean_intensity", "flux_divergence"):
            jac[k] = np.zeros((nif, L + 2))
        out.jac_flux_up = _ptr(jac["flux_up"])
        out.jac_flux_down = _ptr(jac["flux_down"])
        out.jac_mean_intensity = _ptr(jac["mean_intensity"])
        out.jac_flux_divergence = _ptr(jac["flux_divergence"])
    err = C.create_string_buffer(256)
    rc = lib(extended).orc_solve(C.byref(cfg), C.byref(out), err, 256)
    if rc:
        raise OracleError(rc, err.value.decode())
    if compute_temperature_jacobian:
        res["flux_up_temperature_jac"] = jac["flux_up"]
        res["flux_down_temperature_jac"] = jac["flux_down"]
        res["mean_intensity_temperature_jac"] = jac["mean_intensity"]
        res["flux_divergence_temperature_jac"] = jac["flux_divergence"]
    return res


def solve_batch(delta_tau, single_scat_albedo, phase_moments, num_quadrature,
                planck_levels=None, surface_emission=0.0, top_emission=0.0,
                surface_albedo=0.0, solar_flux=0.0, solar_mu=1.0, use_delta_m=False,
                use_diffusion_lower_bc=False, index_from_bottom=False,
                spectrum_scaling=1.0, threads=None, extended=False):
    """Batched columns in wavenumber-major layout, ``[W][L]`` / ``[W][L+1]``.

    ``phase_moments`` is ``[L][M]`` (shared) or ``[W][L][M]``;
    ``surface_emission`` / ``top_emission`` are scalars or ``[W]`` arrays.
    """
    tau = np.ascontiguousarray(delta_tau, dtype=np.float64)
    ssa = np.ascontiguousarray(single_scat_albedo, dtype=np.float64)
    W, L = tau.shape
    mom = np.ascontiguousarray(phase_moments, dtype=np.float64)
    shared = mom.ndim == 2
    M = mom.shape[-1]
    plev = None if planck_levels is None else np.ascontiguousarray(planck_levels, np.float64)
    se_w = None if np.isscalar(surface_emission) else _arr(surface_emission, W)
    te_w = None if np.isscalar(top_emission) else _arr(top_emission, W)
    out = {k: np.zeros((W, L + 1)) for k in ("flux_up", "flux_down", "mean_intensity",
                                               "flux_direct", "flux_divergence")}
    b = _Batch(
        num_wavenumbers=W, num_layers=L, num_quadrature=int(num_quadrature), num_moments=M,
        phase_moments_shared=int(shared), use_delta_m=int(bool(use_delta_m)),
        use_diffusion_lower_bc=int(bool(use_diffusion_lower_bc)),
        index_from_bottom=int(bool(index_from_bottom)),
        surface_albedo=surface_albedo,
        surface_emission=float(surface_emission) if se_w is None else 0.0,
        top_emission=float(top_emission) if te_w is None else 0.0,
        solar_flux=solar_flux, solar_mu=solar_mu, spectrum_scaling=spectrum_scaling,
        delta_tau=_ptr(tau), single_scat_albedo=_ptr(ssa), phase_moments=_ptr(mom),
        planck_levels=_ptr(plev), surface_emission_w=_ptr(se_w), top_emission_w=_ptr(te_w),
        flux_up=_ptr(out["flux_up"]), flux_down=_ptr(out["flux_down"]),
        mean_intensity=_ptr(out["mean_intensity"]), flux_direct=_ptr(out["flux_direct"]),
        flux_divergence=_ptr(out["flux_divergence"]))
    if threads is None:
        threads = min(os.cpu_count() or 1, 16)
    err = C.create_string_buffer(256)
    rc = lib(extended).orc_solve_batch(C.byref(b), int(threads), err, 256)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out


# ---- raw retrieval coefficients (numpy restatement) -------------------------

PLANCK_SINGLE_C1 = 1.19105e-08   # 2 h c^2, scaled to W m^-2 cm (src_cuda/cuda_planck_single.cuh:27)
PLANCK_SINGLE_C2 = 1.43879       # h c / k in cm K (src_cuda/cuda_planck_single.cuh:28)


def planck_single(temperature, wavenumber):
    """Single-wavenumber Planck function B(T, nu) in W m^-2 cm
    (src_cuda/cuda_planck_single.cuh:18-34): c1 nu^3 / (exp(c2 nu / T) - 1),
    0 for T < 1e-5 K.  Broadcasts."""
    T = np.asarray(temperature, dtype=np.float64)
    nu = np.asarray(wavenumber, dtype=np.float64)
    cold = T < 1e-5
    Ts = np.where(cold, 1.0, T)
    with np.errstate(over="ignore"):
        B = PLANCK_SINGLE_C1 * (nu * nu * nu) / np.expm1(PLANCK_SINGLE_C2 * nu / Ts)
    return np.where(cold, 0.0, B)


def raw_optics(absorption_coeff, scattering_coeff, altitude, temperature, wavenumber,
               cloud_optical_depth=None, surface_temperature=-1.0, top_temperature=-1.0):
    """Solver inputs from raw retrieval coefficients, as the reference's
    preprocessRawKernel (src_cuda/cuda_solver.cu:1131-1182) with the boundary
    rules of its inline raw path (src_cuda/cuda_solver.cu:220-241,428-441).

    absorption_coeff / scattering_coeff ``[L+1][W]`` and altitude / temperature
    ``[L+1]`` are bottom of the atmosphere first; cloud_optical_depth ``[L][W]``
    (BOA first) or None.  Layer b (BOA first) spans levels b and b+1:
    tau = dz (ext_top + ext_bot) / 2 (+ cloud), clamped at 0, ext = abs + sca;
    omega = dz (sca_top + sca_bot) / 2 / tau (0 where tau = 0).  Returns
    TOA-first ``delta_tau`` / ``single_scat_albedo`` ``[W][L]``,
    ``planck_levels`` ``[W][L+1]`` and ``surface_emission`` / ``top_emission``
    ``[W]`` (B at surface_temperature / top_temperature when >= 0, else at the
    bottom / top level).
    """
    a = np.asarray(absorption_coeff, dtype=np.float64)
    s = np.asarray(scattering_coeff, dtype=np.float64)
    z = np.asarray(altitude, dtype=np.float64)
    T = np.asarray(temperature, dtype=np.float64)
    nu = np.asarray(wavenumber, dtype=np.float64)
    dz = (z[1:] - z[:-1])[:, None]
    ext = a + s
    tau = dz * (ext[1:] + ext[:-1]) * 0.5
    sdepth = dz * (s[1:] + s[:-1]) * 0.5
    if cloud_optical_depth is not None:
        tau = tau + np.asarray(cloud_optical_depth, dtype=np.float64)
    tau = np.maximum(tau, 0.0)
    ssa = np.zeros_like(tau)
    np.divide(sdepth, tau, out=ssa, where=tau > 0.0)
    L = len(z) - 1
    Ts = surface_temperature if surface_temperature >= 0.0 else T[0]
    Tt = top_temperature if top_temperature >= 0.0 else T[L]
    return {
        "delta_tau": np.ascontiguousarray(tau[::-1].T),
        "single_scat_albedo": np.ascontiguousarray(ssa[::-1].T),
        "planck_levels": np.ascontiguousarray(planck_single(T[::-1][None, :], nu[:, None])),
        "surface_emission": planck_single(Ts, nu),
        "top_emission": planck_single(Tt, nu),
    }
