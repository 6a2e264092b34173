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

    def __init__(self, num_wavenumbers, num_layers, num_quadrature, num_moments, max_bytes=0,
                 adjoint=False, jacobian=False):
        cfg = _lib.BatchConfig(num_wavenumbers=num_wavenumbers, num_layers=num_layers,
                               num_quadrature=num_quadrature, num_moments=num_moments,
                               surface_albedo=0.0, solar_mu=1.0, spectrum_scaling=1.0)
        h = C.c_void_p()
        flags = (_lib.WORKSPACE_ADJOINT if adjoint else 0) | (
            _lib.WORKSPACE_JACOBIAN if jacobian else 0)
        _lib.check(_lib.lib().adrt_workspace_create_ex(C.byref(cfg), int(max_bytes), flags,
                                                       C.byref(h)))
        self.handle = h
        self.adjoint = bool(adjoint)
        self.jacobian = bool(jacobian)
        self.shape = (num_layers, num_quadrature, num_moments)
        b = C.c_size_t()
        ch = C.c_int()
        _lib.check(_lib.lib().adrt_workspace_info(h, C.byref(b), C.byref(ch)))
        self.bytes = b.value
        self.chunk = ch.value
        self.created_for = num_wavenumbers
        # held around every call that enqueues work on this workspace, so that
        # two host threads never interleave their use of its scratch
        self.lock = threading.Lock()

    def generation(self) -> int:
        g = C.c_ulonglong()
        _lib.check(_lib.lib().adrt_workspace_generation(self.handle, C.byref(g)))
        return g.value

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


def get_workspace(num_wavenumbers, num_layers, num_quadrature, num_moments, max_bytes=0,
                  adjoint=False, jacobian=False):
    """The cached workspace for this column shape on the current device and
    stream: work on one stream runs in order, so solves on one stream may share
    scratch; another stream gets a workspace of its own."""
    key = (torch.cuda.current_device(), torch.cuda.current_stream().cuda_stream, num_layers,
           num_quadrature, num_moments, int(max_bytes), bool(adjoint), bool(jacobian))
    with _ws_lock:
        ws = _ws_cache.get(key)
        # a workspace serves any batch size (the library chunks); rebuild it
        # only when its chunk was capped by a smaller earlier batch
        if ws is None or (ws.chunk < num_wavenumbers and ws.chunk == ws.created_for):
            ws = Workspace(num_wavenumbers, num_layers, num_quadrature, num_moments, max_bytes,
                           adjoint, jacobian)
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


def _prepare(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig,
             planck_levels, surface_emission, top_emission):
    W, L = delta_tau.shape
    shared = phase_moments.dim() == 2
    M = phase_moments.shape[-1]
    mshape = (L, M) if shared else (W, L, M)
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
    return W, L, M, shared, cfg, inp


def _outputs_struct(out, W, L):
    return _lib.DeviceOutputs(**{k: _dev_ptr(out[k], k, (W, L + 1)) if out.get(k) is not None
                                 else None for k in OUTPUT_NAMES})


_PREC = {"f64": _lib.PRECISION_F64, "mixed": _lib.PRECISION_MIXED,
         "mixed_reverse": _lib.PRECISION_MIXED_REVERSE}


def solve_batch(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig,
                planck_levels=None, surface_emission=None, top_emission=None,
                precision="f64", out=None, validate=False, max_workspace_bytes=0,
                adjoint_workspace=False):
    """Solve ``W`` independent columns; returns the five ``[W][L+1]`` outputs.

    delta_tau, single_scat_albedo: ``[W][L]``; phase_moments: ``[L][M]`` (shared
    by every wavenumber) or ``[W][L][M]``; planck_levels: ``[W][L+1]`` or None;
    surface_emission / top_emission: ``[W]`` tensors or None (then the scalars
    of ``config`` apply).  With ``adjoint_workspace`` the solve runs on a
    workspace that keeps what ``adjoint_batch`` needs; the returned dict then
    also holds ``"_generation"``, the token that lets the reverse pass reuse it.
    """
    return _solve_batch_ws(delta_tau, single_scat_albedo, phase_moments, config, planck_levels,
                           surface_emission, top_emission, precision, out, validate,
                           max_workspace_bytes, adjoint_workspace)[0]


def _solve_batch_ws(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig,
                    planck_levels=None, surface_emission=None, top_emission=None,
                    precision="f64", out=None, validate=False, max_workspace_bytes=0,
                    adjoint_workspace=False):
    """solve_batch, also returning the workspace the solve ran on."""
    W, L, M, shared, cfg, inp = _prepare(delta_tau, single_scat_albedo, phase_moments, config,
                                         planck_levels, surface_emission, top_emission)
    if out is None:
        out = {k: torch.empty((W, L + 1), dtype=torch.float64, device=delta_tau.device)
               for k in OUTPUT_NAMES}
    o = _outputs_struct(out, W, L)
    ws = get_workspace(W, L, config.num_quadrature, M, max_workspace_bytes, adjoint_workspace)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with ws.lock:
        if validate:
            _lib.check(_lib.lib().adrt_validate_inputs(C.byref(cfg), C.byref(inp), ws.handle,
                                                       stream))
        _lib.check(_lib.lib().adrt_solve_batch(C.byref(cfg), C.byref(inp), C.byref(o),
                                               _PREC[precision], ws.handle, stream))
        if adjoint_workspace:
            out["_generation"] = ws.generation()
    return out, ws


def solve_batch_from_coefficients(absorption_coeff, scattering_coeff, altitude, temperature,
                                  wavenumber, phase_moments, config: BatchConfig,
                                  cloud_optical_depth=None, surface_temperature=-1.0,
                                  top_temperature=-1.0, precision="f64", out=None,
                                  max_workspace_bytes=0):
    """Solve from raw retrieval coefficients (``adrt_solve_batch_from_coefficients``).

    Mirrors ``adrt::cuda::solveBatchFromCoefficients``
    (src_cuda/cuda_solver.cuh:104-179): absorption_coeff / scattering_coeff
    ``[L+1][W]`` and altitude / temperature ``[L+1]`` bottom of the atmosphere
    first, wavenumber ``[W]``, cloud_optical_depth ``[L][W]`` (BOA first) or
    None, phase_moments ``[L][M]`` or ``[W][L][M]`` (TOA first).  Layer optical
    depths use the trapezoid rule, thermal sources the single-wavenumber
    Planck function; a negative surface / top temperature means "use the
    bottom / top level".  Returns the five ``[W][L+1]`` outputs, TOA first.
    ``config.surface_emission``, ``top_emission`` and ``index_from_bottom`` are
    not used.
    """
    W = wavenumber.shape[0]
    L = altitude.shape[0] - 1
    shared = phase_moments.dim() == 2
    M = phase_moments.shape[-1]
    cfg = _lib.BatchConfig(
        num_wavenumbers=W, num_layers=L, num_quadrature=config.num_quadrature, num_moments=M,
        use_delta_m=int(config.use_delta_m),
        use_diffusion_lower_bc=int(config.use_diffusion_lower_bc), index_from_bottom=0,
        surface_albedo=config.surface_albedo, surface_emission=0.0, top_emission=0.0,
        solar_flux=config.solar_flux, solar_mu=config.solar_mu,
        spectrum_scaling=config.spectrum_scaling)
    coef = _lib.DeviceCoefficients(
        absorption_coeff=_dev_ptr(absorption_coeff, "absorption_coeff", (L + 1, W)),
        scattering_coeff=_dev_ptr(scattering_coeff, "scattering_coeff", (L + 1, W)),
        altitude=_dev_ptr(altitude, "altitude", (L + 1,)),
        temperature=_dev_ptr(temperature, "temperature", (L + 1,)),
        wavenumber=_dev_ptr(wavenumber, "wavenumber", (W,)),
        phase_moments=_dev_ptr(phase_moments, "phase_moments", (L, M) if shared else (W, L, M)),
        phase_moments_shared=int(shared),
        cloud_optical_depth=_dev_ptr(cloud_optical_depth, "cloud_optical_depth", (L, W)))
    if out is None:
        out = {k: torch.empty((W, L + 1), dtype=torch.float64, device=wavenumber.device)
               for k in OUTPUT_NAMES}
    o = _outputs_struct(out, W, L)
    # the mixed form below order 9 runs padded to 16 on a reverse-pass workspace
    ws = get_workspace(W, L, config.num_quadrature, M, max_workspace_bytes,
                       precision == "mixed" and config.num_quadrature <= 16)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with ws.lock:
        _lib.check(_lib.lib().adrt_solve_batch_from_coefficients(
            C.byref(cfg), C.byref(coef), float(surface_temperature), float(top_temperature),
            C.byref(o), _PREC[precision], ws.handle, stream))
    return out


GRADIENT_NAMES = ("delta_tau", "single_scat_albedo", "phase_moments", "planck_levels",
                  "surface_emission", "top_emission", "surface_albedo")


def adjoint_batch(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig, outputs,
                  cotangents, planck_levels=None, surface_emission=None, top_emission=None,
                  precision="f64", forward_generation=0, max_workspace_bytes=0,
                  want=GRADIENT_NAMES, check_status=True, workspace=None):
    """Gradients of ``sum(cotangent * output)`` w.r.t. the inputs (``adrt_adjoint_batch``).

    outputs: the dict ``solve_batch`` returned for the same inputs;
    cotangents: dict of ``[W][L+1]`` tensors keyed like the outputs (missing =
    zero).  Returns a dict keyed by ``GRADIENT_NAMES`` (the ones in ``want``):
    ``[W][L]``, moments in their input shape, ``[W][L+1]``, and ``[W]`` for the
    two emissions and the per-column parts of the surface-albedo gradient.
    ``workspace``: the reverse-pass ``Workspace`` of the forward solve (default:
    the cached one of this shape and stream).
    The forward state is reused only if ``forward_generation`` is the token of
    the last solve on that workspace; otherwise it is recomputed.
    """
    W, L, M, shared, cfg, inp = _prepare(delta_tau, single_scat_albedo, phase_moments, config,
                                         planck_levels, surface_emission, top_emission)
    dev = delta_tau.device
    shapes = {"delta_tau": (W, L), "single_scat_albedo": (W, L),
              "phase_moments": (L, M) if shared else (W, L, M), "planck_levels": (W, L + 1),
              "surface_emission": (W,), "top_emission": (W,), "surface_albedo": (W,)}
    grads = {k: torch.empty(shapes[k], dtype=torch.float64, device=dev) for k in want}
    g = _lib.DeviceGradients(**{k: C.c_void_p(v.data_ptr()) for k, v in grads.items()})
    cot = {}
    for k in OUTPUT_NAMES:
        c = cotangents.get(k)
        if c is not None:
            c = c.to(dtype=torch.float64).contiguous()
            cot[k] = c
    ct = _lib.DeviceCotangents(**{k: _dev_ptr(v, "cotangent " + k, (W, L + 1))
                                  for k, v in cot.items()})
    o = _outputs_struct(outputs, W, L)
    ws = workspace if workspace is not None else get_workspace(
        W, L, config.num_quadrature, M, max_workspace_bytes, True)
    if not ws.adjoint or ws.shape != (L, config.num_quadrature, M):
        raise ValueError("workspace is not a reverse-pass workspace of this column shape")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with ws.lock:
        _lib.check(_lib.lib().adrt_adjoint_batch(C.byref(cfg), C.byref(inp), C.byref(o),
                                                 C.byref(ct), C.byref(g), _PREC[precision],
                                                 int(forward_generation), ws.handle, stream))
        if check_status:
            _lib.check(_lib.lib().adrt_adjoint_status(ws.handle, stream))
    return grads


REVERSE_MAX_QUADRATURE = 32  # every order the forward takes


JACOBIAN_NAMES = ("flux_up", "flux_down", "mean_intensity", "flux_divergence")


def source_jacobians(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig,
                     divergence=True, max_workspace_bytes=0):
    """Derivatives of the outputs with respect to the thermal sources
    (``adrt_source_jacobian_batch``): a dict of ``[W][L+3][L+1]`` tensors,
    ``J[w, u, i] = d output(w, interface i) / d source u`` with u = 0..L the
    Planck value of level u, L+1 the surface emission and L+2 the top emission
    (the part of the reference's computeTemperatureJacobian, src/solver.cpp:
    117-341, before the dB/dT chain rule).  Each column's doubling runs once;
    any quadrature order."""
    W, L, M, shared, cfg, inp = _prepare(delta_tau, single_scat_albedo, phase_moments, config,
                                         None, None, None)
    dev = delta_tau.device
    names = JACOBIAN_NAMES if divergence else JACOBIAN_NAMES[:3]
    jac = {k: torch.empty((W, L + 3, L + 1), dtype=torch.float64, device=dev) for k in names}
    js = _lib.DeviceJacobians(**{k: C.c_void_p(v.data_ptr()) for k, v in jac.items()})
    ws = get_workspace(W, L, config.num_quadrature, M, max_workspace_bytes, jacobian=True)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with ws.lock:
        _lib.check(_lib.lib().adrt_source_jacobian_batch(C.byref(cfg), C.byref(inp), C.byref(js),
                                                         ws.handle, stream))
    return jac


def planck_source_vjp(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig,
                      cotangents, max_columns=1 << 14, max_workspace_bytes=0):
    """Gradients of ``sum(cotangent * output)`` w.r.t. the Planck levels and the
    surface and top emission, from the source Jacobians (any quadrature order).

    R and T do not depend on the thermal sources and every output is affine in
    them, so these gradients are the cotangents contracted with
    ``source_jacobians`` (the linear recursion the reference's
    computeTemperatureJacobian runs in forward mode, src/solver.cpp:117-341),
    in batches of at most ``max_columns`` columns.  Returns ``[W][L+1]``,
    ``[W]``, ``[W]``.
    """
    W, L = delta_tau.shape
    dev = delta_tau.device
    gB = torch.zeros((W, L + 1), dtype=torch.float64, device=dev)
    gS = torch.zeros(W, dtype=torch.float64, device=dev)
    gT = torch.zeros(W, dtype=torch.float64, device=dev)
    div = cotangents.get("flux_divergence") is not None
    for w0 in range(0, W, max_columns):
        w1 = min(W, w0 + max_columns)
        mom = phase_moments if phase_moments.dim() == 2 else phase_moments[w0:w1].contiguous()
        J = source_jacobians(delta_tau[w0:w1].contiguous(),
                             single_scat_albedo[w0:w1].contiguous(), mom, config,
                             divergence=div, max_workspace_bytes=max_workspace_bytes)
        acc = torch.zeros((w1 - w0, L + 3), dtype=torch.float64, device=dev)
        for k in JACOBIAN_NAMES:
            c = cotangents.get(k)
            if c is None or k not in J:
                continue
            acc += torch.bmm(J[k], c[w0:w1].to(torch.float64).unsqueeze(-1)).squeeze(-1)
        gB[w0:w1] = acc[:, :L + 1]
        gS[w0:w1] = acc[:, L + 1]
        gT[w0:w1] = acc[:, L + 2]
    return gB, gS, gT


class _SolveFunction(torch.autograd.Function):
    """Forward solve with a hand-written reverse pass (adrt_adjoint_batch)."""

    @staticmethod
    def forward(ctx, config, precision, max_ws, delta_tau, single_scat_albedo, phase_moments,
                planck_levels, surface_emission, top_emission, surface_albedo):
        if surface_albedo is not None:
            config = BatchConfig(**{**config.__dict__, "surface_albedo": float(surface_albedo)})
        out, ws = _solve_batch_ws(delta_tau, single_scat_albedo, phase_moments, config,
                                  planck_levels, surface_emission, top_emission,
                                  precision=precision, max_workspace_bytes=max_ws,
                                  adjoint_workspace=True)
        gen = out.pop("_generation")
        # the workspace that holds this forward's state: a later solve of
        # another shape or size may replace the cached one before backward runs
        ctx.ws = ws
        ctx.config, ctx.precision, ctx.max_ws, ctx.generation = config, precision, max_ws, gen
        ctx.has = (planck_levels is not None, surface_emission is not None,
                   top_emission is not None, surface_albedo is not None)
        outs = tuple(out[k] for k in OUTPUT_NAMES)
        ctx.save_for_backward(delta_tau, single_scat_albedo, phase_moments, planck_levels,
                              surface_emission, top_emission, *outs)
        return outs

    @staticmethod
    def backward(ctx, *gouts):
        tau, ssa, mom, B, Bs, Bt, *outs = ctx.saved_tensors
        out = dict(zip(OUTPUT_NAMES, outs))
        cot = {k: g for k, g in zip(OUTPUT_NAMES, gouts) if g is not None}
        hasB, hasBs, hasBt, hasA = ctx.has
        want = ["delta_tau", "single_scat_albedo", "phase_moments"]
        if hasB:
            want.append("planck_levels")
        if hasBs:
            want.append("surface_emission")
        if hasBt:
            want.append("top_emission")
        if hasA:
            want.append("surface_albedo")
        g = adjoint_batch(tau, ssa, mom, ctx.config, out, cot, B, Bs, Bt,
                          precision=ctx.precision, forward_generation=ctx.generation,
                          max_workspace_bytes=ctx.max_ws, want=tuple(want), workspace=ctx.ws)
        ctx.ws = None
        gA = g["surface_albedo"].sum() if hasA else None
        return (None, None, None, g["delta_tau"], g["single_scat_albedo"], g["phase_moments"],
                g.get("planck_levels"), g.get("surface_emission"), g.get("top_emission"), gA)


def solve(delta_tau, single_scat_albedo, phase_moments, config: BatchConfig, planck_levels=None,
          surface_emission=None, top_emission=None, surface_albedo=None, precision="f64",
          max_workspace_bytes=0):
    """Differentiable batched solve (``torch.autograd``); returns the five outputs as a dict.

    Gradients flow to delta_tau, single_scat_albedo, phase_moments,
    planck_levels, surface_emission, top_emission and ``surface_albedo`` (a
    0-dim tensor that replaces ``config.surface_albedo``) through the
    hand-written reverse pass (every quadrature order; above 16 in FP64
    only).  flux_direct,
    F0 mu0 exp(-tau_cum / mu0) with tau_cum the cumulative delta-M-scaled depth,
    carries gradients to delta_tau, single_scat_albedo and (through delta-M) the
    phase moments like the other outputs.
    """
    outs = _SolveFunction.apply(config, precision, max_workspace_bytes, delta_tau,
                                single_scat_albedo, phase_moments, planck_levels,
                                surface_emission, top_emission, surface_albedo)
    return dict(zip(OUTPUT_NAMES, outs))


def _temperature_value(t):
    if isinstance(t, torch.Tensor):
        if t.numel() != 1:
            raise ValueError("a boundary temperature tensor must hold one value")
        return float(t.detach().item())
    return float(t)


def _coefficients_struct(a, s, z, T, nu, cloud, L, W):
    return _lib.DeviceCoefficients(
        absorption_coeff=_dev_ptr(a, "absorption_coeff", (L + 1, W)),
        scattering_coeff=_dev_ptr(s, "scattering_coeff", (L + 1, W)),
        altitude=_dev_ptr(z, "altitude", (L + 1,)),
        temperature=_dev_ptr(T, "temperature", (L + 1,)),
        wavenumber=_dev_ptr(nu, "wavenumber", (W,)),
        cloud_optical_depth=_dev_ptr(cloud, "cloud_optical_depth", (L, W)))


RAW_OPTICS_NAMES = ("delta_tau", "single_scat_albedo", "planck_levels", "surface_emission",
                    "top_emission")


class _RawOpticsFunction(torch.autograd.Function):
    """adrt_raw_optics forward, adrt_raw_optics_adjoint backward."""

    @staticmethod
    def forward(ctx, a, s, z, T, nu, cloud, Ts, Tt):
        W, L = nu.shape[0], z.shape[0] - 1
        Tsv, Ttv = _temperature_value(Ts), _temperature_value(Tt)
        coef = _coefficients_struct(a, s, z, T, nu, cloud, L, W)
        dev = nu.device
        e = lambda *shape: torch.empty(shape, dtype=torch.float64, device=dev)
        outs = (e(W, L), e(W, L), e(W, L + 1), e(W), e(W))
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().adrt_raw_optics(W, L, C.byref(coef), Tsv, Ttv,
                                              *(C.c_void_p(o.data_ptr()) for o in outs),
                                              stream))
        ctx.save_for_backward(a, s, z, T, nu, cloud)
        ctx.temps = (Tsv, Ttv, Ts.shape if isinstance(Ts, torch.Tensor) else None,
                     Tt.shape if isinstance(Tt, torch.Tensor) else None)
        return outs

    @staticmethod
    def backward(ctx, *gouts):
        a, s, z, T, nu, cloud = ctx.saved_tensors
        W, L = nu.shape[0], z.shape[0] - 1
        Tsv, Ttv, ts_shape, tt_shape = ctx.temps
        need = ctx.needs_input_grad
        dev = nu.device
        e = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=dev)
        gr = {"absorption_coeff": e(L + 1, W) if need[0] else None,
              "scattering_coeff": e(L + 1, W) if need[1] else None,
              "cloud_optical_depth": e(L, W) if need[5] and cloud is not None else None,
              "temperature": e(L + 1) if need[3] else None,
              "surface_temperature": e(1) if need[6] and ts_shape is not None else None,
              "top_temperature": e(1) if need[7] and tt_shape is not None else None}
        cot = [None if g is None else g.to(torch.float64).contiguous() for g in gouts]
        ct = _lib.RawOpticsCotangents(**{k: None if g is None else C.c_void_p(g.data_ptr())
                                         for k, g in zip(RAW_OPTICS_NAMES, cot)})
        g = _lib.RawGradients(**{k: None if v is None else C.c_void_p(v.data_ptr())
                                 for k, v in gr.items()})
        coef = _coefficients_struct(a, s, z, T, nu, cloud, L, W)
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().adrt_raw_optics_adjoint(W, L, C.byref(coef), Tsv, Ttv,
                                                      C.byref(ct), C.byref(g), stream))
        ts = gr["surface_temperature"]
        tt = gr["top_temperature"]
        return (gr["absorption_coeff"], gr["scattering_coeff"], None, gr["temperature"], None,
                gr["cloud_optical_depth"], None if ts is None else ts.reshape(ts_shape),
                None if tt is None else tt.reshape(tt_shape))


def raw_optics(absorption_coeff, scattering_coeff, altitude, temperature, wavenumber,
               cloud_optical_depth=None, surface_temperature=-1.0, top_temperature=-1.0):
    """Optics of the raw-coefficient entry (``adrt_raw_optics``), differentiable.

    Inputs as ``solve_batch_from_coefficients``; surface_temperature and
    top_temperature may be 0-dim tensors to get their gradients (a negative
    value means "use the bottom / top level", as there).  Returns a dict of
    delta_tau and single_scat_albedo ``[W][L]``, planck_levels ``[W][L+1]``,
    surface_emission and top_emission ``[W]``, TOA first: the arrays the
    reference's preprocessRawKernel (src_cuda/cuda_solver.cu:1131-1182)
    prepares.  Gradients flow to the coefficients, the cloud depths and the
    temperatures; altitudes and wavenumbers get none, and a layer whose optical
    depth is clamped at zero passes nothing back.
    """
    outs = _RawOpticsFunction.apply(absorption_coeff, scattering_coeff, altitude, temperature,
                                    wavenumber, cloud_optical_depth, surface_temperature,
                                    top_temperature)
    r = dict(zip(RAW_OPTICS_NAMES, outs))
    grad = lambda *xs: any(isinstance(x, torch.Tensor) and x.requires_grad for x in xs)
    # each output depends on one group of inputs only: keep the others off the graph
    if not grad(absorption_coeff, scattering_coeff, cloud_optical_depth):
        for k in RAW_OPTICS_NAMES[:2]:
            r[k] = r[k].detach()
    if not grad(temperature, surface_temperature, top_temperature):
        for k in RAW_OPTICS_NAMES[2:]:
            r[k] = r[k].detach()
    return r


def solve_from_coefficients(absorption_coeff, scattering_coeff, altitude, temperature,
                            wavenumber, phase_moments, config: BatchConfig,
                            cloud_optical_depth=None, surface_temperature=-1.0,
                            top_temperature=-1.0, surface_albedo=None, precision="f64",
                            max_workspace_bytes=0):
    """Differentiable solve from raw retrieval coefficients: ``raw_optics``
    chained with ``solve``.  The forward result equals
    ``solve_batch_from_coefficients`` on the same inputs; gradients flow to the
    coefficients, cloud depths, temperatures, phase moments and surface_albedo."""
    r = raw_optics(absorption_coeff, scattering_coeff, altitude, temperature, wavenumber,
                   cloud_optical_depth, surface_temperature, top_temperature)
    return solve(r["delta_tau"], r["single_scat_albedo"], phase_moments, config,
                 planck_levels=r["planck_levels"], surface_emission=r["surface_emission"],
                 top_emission=r["top_emission"], surface_albedo=surface_albedo,
                 precision=precision, max_workspace_bytes=max_workspace_bytes)


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


def _planck_levels_raw(wn_lo, wn_hi, temperature, derivative):
    W = wn_lo.shape[0]
    shared = temperature.dim() == 1
    nlev = temperature.shape[-1]
    out = torch.empty((W, nlev), dtype=torch.float64, device=wn_lo.device)
    ws = get_workspace(1, 1, 2, 1)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with ws.lock:
        _lib.check(_lib.lib().adrt_planck_levels(
            _dev_ptr(wn_lo, "wn_lo", (W,)), _dev_ptr(wn_hi, "wn_hi", (W,)),
            _dev_ptr(temperature.contiguous(), "temperature"), int(shared), W, nlev,
            int(bool(derivative)), _dev_ptr(out, "out"), ws.handle, stream))
    return out


class _PlanckFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, wn_lo, wn_hi, temperature):
        ctx.save_for_backward(wn_lo, wn_hi, temperature)
        return _planck_levels_raw(wn_lo, wn_hi, temperature, False)

    @staticmethod
    def backward(ctx, g):
        lo, hi, T = ctx.saved_tensors
        dB = _planck_levels_raw(lo, hi, T.detach(), True)
        gT = g * dB
        if T.dim() == 1:
            gT = gT.sum(dim=0)
        return None, None, gT


def planck_levels(wn_lo, wn_hi, temperature, derivative=False):
    """Band Planck (or dB/dT) at every level on the device.

    wn_lo, wn_hi: ``[W]`` band edges; temperature: ``[L+1]`` (shared) or
    ``[W][L+1]``.  Returns ``[W][L+1]`` (src/planck.cpp:35-177).  Differentiable
    in ``temperature`` (chain rule with the device dB/dT).
    """
    if derivative:
        return _planck_levels_raw(wn_lo, wn_hi, temperature, True)
    return _PlanckFunction.apply(wn_lo, wn_hi, temperature)
