"""ctypes binding of the C-ABI in include/adrt_hip.h.

The HIP library ``libadrt_hip.so`` is built in-tree (``build.py``).  Every
product entry point goes through it; there is no CPU fallback: a missing or
unloadable library raises ``AdrtLibraryError``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadrt_hip.so")

ADRT_OK = 0
PRECISION_F64 = 0
PRECISION_MIXED = 1
PRECISION_MIXED_REVERSE = 2
WORKSPACE_ADJOINT = 1
WORKSPACE_JACOBIAN = 2
PROFILE_KINDS = 7
NOALLOC_SCRATCH_BYTES = 64 << 20  # ADRT_NOALLOC_SCRATCH_BYTES

_dp = C.POINTER(C.c_double)


class AdrtLibraryError(RuntimeError):
    pass


class AdrtError(ValueError):
    """A non-zero status from the library (validation failures are 1..17)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class BatchConfig(C.Structure):
    _fields_ = [
        ("num_wavenumbers", C.c_int), ("num_layers", C.c_int), ("num_quadrature", C.c_int),
        ("num_moments", C.c_int), ("use_delta_m", C.c_int), ("use_diffusion_lower_bc", C.c_int),
        ("index_from_bottom", C.c_int),
        ("surface_albedo", C.c_double), ("surface_emission", C.c_double),
        ("top_emission", C.c_double), ("solar_flux", C.c_double), ("solar_mu", C.c_double),
        ("spectrum_scaling", C.c_double),
    ]


class DeviceInputs(C.Structure):
    _fields_ = [
        ("delta_tau", C.c_void_p), ("single_scat_albedo", C.c_void_p),
        ("phase_moments", C.c_void_p), ("phase_moments_shared", C.c_int),
        ("planck_levels", C.c_void_p), ("surface_emission", C.c_void_p),
        ("top_emission", C.c_void_p),
    ]


class DeviceOutputs(C.Structure):
    _fields_ = [
        ("flux_up", C.c_void_p), ("flux_down", C.c_void_p), ("mean_intensity", C.c_void_p),
        ("flux_direct", C.c_void_p), ("flux_divergence", C.c_void_p),
    ]


class DeviceCotangents(C.Structure):
    _fields_ = [
        ("flux_up", C.c_void_p), ("flux_down", C.c_void_p), ("mean_intensity", C.c_void_p),
        ("flux_direct", C.c_void_p), ("flux_divergence", C.c_void_p),
    ]


class DeviceGradients(C.Structure):
    _fields_ = [
        ("delta_tau", C.c_void_p), ("single_scat_albedo", C.c_void_p),
        ("phase_moments", C.c_void_p), ("planck_levels", C.c_void_p),
        ("surface_emission", C.c_void_p), ("top_emission", C.c_void_p),
        ("surface_albedo", C.c_void_p),
    ]


class DeviceJacobians(C.Structure):
    _fields_ = [
        ("flux_up", C.c_void_p), ("flux_down", C.c_void_p), ("mean_intensity", C.c_void_p),
        ("flux_divergence", C.c_void_p),
    ]


class DeviceCoefficients(C.Structure):
    _fields_ = [
        ("absorption_coeff", C.c_void_p), ("scattering_coeff", C.c_void_p),
        ("altitude", C.c_void_p), ("temperature", C.c_void_p), ("wavenumber", C.c_void_p),
        ("phase_moments", C.c_void_p), ("phase_moments_shared", C.c_int),
        ("cloud_optical_depth", C.c_void_p),
    ]


class RawOpticsCotangents(C.Structure):
    _fields_ = [
        ("delta_tau", C.c_void_p), ("single_scat_albedo", C.c_void_p),
        ("planck_levels", C.c_void_p), ("surface_emission", C.c_void_p),
        ("top_emission", C.c_void_p),
    ]


class RawGradients(C.Structure):
    _fields_ = [
        ("absorption_coeff", C.c_void_p), ("scattering_coeff", C.c_void_p),
        ("cloud_optical_depth", C.c_void_p), ("temperature", C.c_void_p),
        ("surface_temperature", C.c_void_p), ("top_temperature", C.c_void_p),
    ]


# Every symbol include/adrt_hip.h declares, with its ctypes signature.
SIGNATURES = {
    "adrt_workspace_create": (C.c_int, [C.POINTER(BatchConfig), C.c_size_t, C.POINTER(C.c_void_p)]),
    "adrt_workspace_create_ex": (C.c_int, [C.POINTER(BatchConfig), C.c_size_t, C.c_int,
                                           C.POINTER(C.c_void_p)]),
    "adrt_workspace_destroy": (C.c_int, [C.c_void_p]),
    "adrt_workspace_generation": (C.c_int, [C.c_void_p, C.POINTER(C.c_ulonglong)]),
    "adrt_adjoint_batch": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(DeviceInputs),
                                     C.POINTER(DeviceOutputs), C.POINTER(DeviceCotangents),
                                     C.POINTER(DeviceGradients), C.c_int, C.c_ulonglong,
                                     C.c_void_p, C.c_void_p]),
    "adrt_adjoint_status": (C.c_int, [C.c_void_p, C.c_void_p]),
    "adrt_workspace_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int)]),
    "adrt_solve_batch": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(DeviceInputs),
                                   C.POINTER(DeviceOutputs), C.c_int, C.c_void_p, C.c_void_p]),
    "adrt_solve_batch_noalloc": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(DeviceInputs),
                                           C.POINTER(DeviceOutputs), C.c_void_p]),
    "adrt_source_jacobian_batch": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(DeviceInputs),
                                             C.POINTER(DeviceJacobians), C.c_void_p, C.c_void_p]),
    "adrt_solve_batch_from_coefficients": (C.c_int, [C.POINTER(BatchConfig),
                                                     C.POINTER(DeviceCoefficients), C.c_double,
                                                     C.c_double, C.POINTER(DeviceOutputs), C.c_int,
                                                     C.c_void_p, C.c_void_p]),
    "adrt_raw_optics": (C.c_int, [C.c_int, C.c_int, C.POINTER(DeviceCoefficients), C.c_double,
                                  C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "adrt_raw_optics_adjoint": (C.c_int, [C.c_int, C.c_int, C.POINTER(DeviceCoefficients),
                                          C.c_double, C.c_double, C.POINTER(RawOpticsCotangents),
                                          C.POINTER(RawGradients), C.c_void_p]),
    "adrt_validate_inputs": (C.c_int, [C.POINTER(BatchConfig), C.POINTER(DeviceInputs),
                                       C.c_void_p, C.c_void_p]),
    "adrt_planck_levels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "adrt_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "adrt_profile_read": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int)]),
    "adrt_gauss_legendre": (C.c_int, [C.c_int, _dp, _dp]),
    "adrt_planck": (C.c_int, [C.c_double, C.c_double, C.c_double, _dp]),
    "adrt_planck_deriv": (C.c_int, [C.c_double, C.c_double, C.c_double, _dp]),
    "adrt_status_string": (C.c_char_p, [C.c_int]),
    "adrt_version": (C.c_char_p, []),
}

_lib = None


def lib():
    """Load libadrt_hip.so once; raise AdrtLibraryError if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AdrtLibraryError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - environment problem
            raise AdrtLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != ADRT_OK:
        msg = lib().adrt_status_string(status).decode()
        raise AdrtError(status, msg)
