"""Batched adding-doubling solve on torch device tensors through the C-ABI.

``solve_batch`` is the Python face of ``adrt_solve_batch`` (include/adrt_hip.h):
inputs and outputs are FP64 device tensors in wavenumber-major layout, the
work is enqueued on torch's current HIP stream, and a workspace is cached per
column shape.  Field names follow the reference's batched API
(``adrt::cuda::BatchConfig`` / ``DeviceData``, src_cuda/cuda_solver.cuh:18-68).
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import torch

from . import _lib

OUTPUT_NAMES = ("flux_up", "flux_down", "mean_intensity", "flux_direct", "flux_divergence")


@dataclass(frozen=True)
class BatchConfig:
    """Scalars shared by all columns (src_cuda/cuda_solver.cuh:18-40)."""
    num_quadrature: int = 8
    use_delta_m: bool = False
    use_diffusion_lower_bc: bool = False
    index_from_bottom: bool = False
    surface_albedo: float = 0.0
    surface_emission: float = 0.0
    top_emission: float = 0.0
    solar_flux: float = 0.0
    solar_mu: float = 1.0
    spectrum_scaling: float = 1.0


class Workspace:
    """Owns an ``adrt_workspace`` (device scratch for one column shape)."""

    def __init__(self, num_wavenumbers, num_layers, num_quadrature, num_moments, max_bytes=0):
        cfg = _lib.BatchConfig(num_wavenumbers=num_wavenumbers, num_layers=num_layers,
                               num_quadrature=num_quadrature, num_moments=num_moments,
                               surface_albedo=0.0, solar_mu=1.0, spectrum_scaling=1.0)
        h = C.c_void_p()
        _lib.check(_lib.lib().adrt_workspace_create(C.byref(cfg), int(max_bytes), C.byref(h)))
        self.handle = h
        self.shape = (num_layers, num_quadrature, num_moments)
        b = C.c_size_t()
        ch = C.c_int()
        _lib.check(_lib.lib().adrt_workspace_info(h, C.byref(b), C.byref(ch)))
        self.bytes = b.value
        self.chunk = ch.value
        self.created_for = num_wavenumbers

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().adrt_workspace_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self.handle = None


_ws_cache: dict = {}
_ws_lock = threading.Lock()


def get_workspace(num_wavenumbers, num_layers, num_quadrature, num_moments, max_bytes=0):
    key = (torch.cuda.current_device(), num_layers, num_quadrature, num_moments, int(max_bytes))
    with _ws_lock:
        ws = _ws_cache.get(key)
        # a workspace serves any batch size (the library chunks); rebuild it
        # only when its chunk was capped by a smaller earlier batch
        if ws is None or (ws.chunk < num_wavenumbers and ws.chunk == ws.created_for):
            ws = Workspace(num_wavenumbers, num_layers, num_quadrature, num_moments, max_bytes)
            _ws_cache[key] = ws
        return ws


def _dev_ptr(t, name, shape=None):
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous float64 device tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return C.c_void_p(t.data_ptr())


def solve_batch(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig,
                planck_levels=None, surface_emission=None, top_emission=None,
                precision="f64", out=None, validate=False, max_workspace_bytes=0):
    """Solve ``W`` independent columns; returns the five ``[W][L+1]`` outputs.

    delta_tau, single_scat_albedo: ``[W][L]``; phase_moments: ``[L][M]`` (shared
    by every wavenumber) or ``[W][L][M]``; planck_levels: ``[W][L+1]`` or None;
    surface_emission / top_emission: ``[W]`` tensors or None (then the scalars
    of ``config`` apply).
    """
    W, L = delta_tau.shape
    shared = phase_moments.dim() == 2
    M = phase_moments.shape[-1]
    mshape = (L, M) if shared else (W, L, M)
    prec = {"f64": _lib.PRECISION_F64, "mixed": _lib.PRECISION_MIXED}[precision]
    cfg = _lib.BatchConfig(
        num_wavenumbers=W, num_layers=L, num_quadrature=config.num_quadrature, num_moments=M,
        use_delta_m=int(config.use_delta_m), use_diffusion_lower_bc=int(config.use_diffusion_lower_bc),
        index_from_bottom=int(config.index_from_bottom), surface_albedo=config.surface_albedo,
        surface_emission=config.surface_emission, top_emission=config.top_emission,
        solar_flux=config.solar_flux, solar_mu=config.solar_mu,
        spectrum_scaling=config.spectrum_scaling)
    inp = _lib.DeviceInputs(
        delta_tau=_dev_ptr(delta_tau, "delta_tau", (W, L)),
        single_scat_albedo=_dev_ptr(single_scat_albedo, "single_scat_albedo", (W, L)),
        phase_moments=_dev_ptr(phase_moments, "phase_moments", mshape),
        phase_moments_shared=int(shared),
        planck_levels=_dev_ptr(planck_levels, "planck_levels", (W, L + 1)),
        surface_emission=_dev_ptr(surface_emission, "surface_emission", (W,)),
        top_emission=_dev_ptr(top_emission, "top_emission", (W,)))
    if out is None:
        out = {k: torch.empty((W, L + 1), dtype=torch.float64, device=delta_tau.device)
               for k in OUTPUT_NAMES}
    o = _lib.DeviceOutputs(**{k: _dev_ptr(out[k], k, (W, L + 1)) if out.get(k) is not None
                              else None for k in OUTPUT_NAMES})
    ws = get_workspace(W, L, config.num_quadrature, M, max_workspace_bytes)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if validate:
        _lib.check(_lib.lib().adrt_validate_inputs(C.byref(cfg), C.byref(inp), ws.handle, stream))
    _lib.check(_lib.lib().adrt_solve_batch(C.byref(cfg), C.byref(inp), C.byref(o), prec,
                                           ws.handle, stream))
    return out


def gauss_legendre(n: int):
    import numpy as np
    mu = np.zeros(n)
    wt = np.zeros(n)
    _lib.check(_lib.lib().adrt_gauss_legendre(n, mu.ctypes.data_as(_lib._dp),
                                              wt.ctypes.data_as(_lib._dp)))
    return mu, wt


def planck(lo: float, hi: float, T: float) -> float:
    out = C.c_double()
    _lib.check(_lib.lib().adrt_planck(lo, hi, T, C.byref(out)))
    return out.value


def planck_deriv(lo: float, hi: float, T: float) -> float:
    out = C.c_double()
    _lib.check(_lib.lib().adrt_planck_deriv(lo, hi, T, C.byref(out)))
    return out.value


def planck_levels(wn_lo, wn_hi, temperature, derivative=False):
    """Band Planck (or dB/dT) at every level on the device.

    wn_lo, wn_hi: ``[W]`` band edges; temperature: ``[L+1]`` (shared) or
    ``[W][L+1]``.  Returns ``[W][L+1]`` (src/planck.cpp:35-177).
    """
    W = wn_lo.shape[0]
    shared = temperature.dim() == 1
    nlev = temperature.shape[-1]
    out = torch.empty((W, nlev), dtype=torch.float64, device=wn_lo.device)
    ws = get_workspace(1, 1, 2, 1)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().adrt_planck_levels(
        _dev_ptr(wn_lo, "wn_lo", (W,)), _dev_ptr(wn_hi, "wn_hi", (W,)),
        _dev_ptr(temperature, "temperature"), int(shared), W, nlev, int(bool(derivative)),
        _dev_ptr(out, "out"), ws.handle, stream))
    return out
