"""Synthetic column sets of the five benchmark configurations.

Every configuration is generated on the host with numpy
``Generator(PCG64(20261203 + k))`` (k = 1..5), draw order tau, omega,
phase parameters, temperature noise, in float64.  Choices the configuration
text leaves open are fixed here (and recorded in the design notes):

* thermal bands: W equal bins spanning 500-2500 cm^-1 (bin w = [lo_w, hi_w]);
* thermal mode follows ADConfig semantics: the TOA downwelling is B(T[0])
  (top_temperature < 0) and the surface emits B(surface_temperature) or
  B(T[L]);
* config 3: temperatures 180 K (top) -> 300 K (bottom), solar mu0 = 0.5;
* config 4: surface albedo 0.1 (so the albedo gradient is non-trivial);
* config 5: tau log-uniform [1e-3, 10] and omega ~ U(0, 1) (as config 2).

``make(k, W=None)`` returns a ``Columns`` with host arrays; ``W`` overrides
the number of wavenumbers (parity tests run reduced W of the same config).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SEED0 = 20261203


@dataclass
class Columns:
    name: str
    W: int
    L: int
    N: int
    delta_tau: np.ndarray             # [W][L]
    single_scat_albedo: np.ndarray    # [W][L]
    phase_moments: np.ndarray         # [L][M] or [W][L][M]
    temperature: np.ndarray | None    # [W][L+1] or [L+1] (shared) or None
    wn_lo: np.ndarray | None          # [W]
    wn_hi: np.ndarray | None          # [W]
    surface_temperature: float = -1.0
    surface_albedo: float = 0.0
    solar_flux: float = 0.0
    solar_mu: float = 1.0
    use_delta_m: bool = False
    use_diffusion_lower_bc: bool = False
    loss_weights: np.ndarray | None = None  # [W] for the gradient workload
    extra: dict = field(default_factory=dict)

    @property
    def thermal(self) -> bool:
        return self.temperature is not None


def _bands(W):
    edges = np.linspace(500.0, 2500.0, W + 1)
    return edges[:-1].copy(), edges[1:].copy()


def hg_moments(g, M):
    """chi_l = g^l (ADConfig::setHenyeyGreenstein, src/adding_doubling.hpp:181-193)."""
    g = np.asarray(g, dtype=np.float64)
    return g[..., None] ** np.arange(M)


def make(k: int, W: int | None = None) -> Columns:
    rng = np.random.Generator(np.random.PCG64(SEED0 + k))
    if k == 1:
        W = W or 1000
        L, N = 50, 8
        tau = rng.uniform(0.01, 0.5, (W, L))
        ssa = np.full((W, L), 0.9)
        mom = np.tile(hg_moments(0.7, 16), (L, 1))
        return Columns("c1_f64_fwd_N8", W, L, N, tau, ssa, mom, None, None, None,
                       surface_albedo=0.1, solar_flux=1.0, solar_mu=0.5)
    if k == 2:
        W = W or 16384
        L, N, M = 100, 16, 32
        tau = 10.0 ** rng.uniform(-3.0, 1.0, (W, L))
        ssa = rng.uniform(0.0, 1.0, (W, L))
        g1 = rng.uniform(0.5, 0.9, (W, L))
        g2 = rng.uniform(-0.5, -0.1, (W, L))
        f = rng.uniform(0.7, 1.0, (W, L))
        mom = f[..., None] * hg_moments(g1, M) + (1.0 - f[..., None]) * hg_moments(g2, M)
        T = np.linspace(180.0, 300.0, L + 1)
        lo, hi = _bands(W)
        return Columns("c2_f64_fwd_N16", W, L, N, tau, ssa, mom, T, lo, hi,
                       surface_temperature=310.0, surface_albedo=0.2, solar_flux=1.0,
                       solar_mu=0.6, use_delta_m=True)
    if k == 3:
        W = W or 8192
        L, N, M = 80, 32, 64
        tau = 10.0 ** rng.uniform(-2.0, 2.0, (W, L))
        ssa = 1.0 - 10.0 ** rng.uniform(-6.0, -1.0, (W, L))
        g = rng.uniform(0.6, 0.95, (W, L))
        mom = hg_moments(g, M)
        T = np.linspace(180.0, 300.0, L + 1)
        lo, hi = _bands(W)
        return Columns("c3_mixed_fwd_N32", W, L, N, tau, ssa, mom, T, lo, hi,
                       solar_flux=1.0, solar_mu=0.5, use_diffusion_lower_bc=True)
    if k == 4:
        W = W or 4096
        L, N, M = 60, 8, 16
        tau = 10.0 ** rng.uniform(-3.0, np.log10(5.0), (W, L))
        ssa = rng.uniform(0.0, 0.99, (W, L))
        mom = np.zeros((L, M))
        mom[:, 0] = 1.0
        mom[:, 2] = 0.1  # Rayleigh (ADConfig::setRayleigh)
        T = np.linspace(320.0, 200.0, L + 1)[::-1].copy()  # 200 K at TOA .. 320 K at the bottom
        T = T + rng.normal(0.0, 5.0, L + 1)
        lo, hi = _bands(W)
        u = rng.uniform(0.0, 1.0, W)
        return Columns("c4_f64_fwd_adj_N8", W, L, N, tau, ssa, mom, T, lo, hi,
                       surface_albedo=0.1, loss_weights=u)
    if k == 5:
        W = W or 65536
        L, N, M = 100, 16, 32
        tau = 10.0 ** rng.uniform(-3.0, 1.0, (W, L))
        ssa = rng.uniform(0.0, 1.0, (W, L))
        g = rng.uniform(0.5, 0.85, L)
        r = rng.uniform(0.0, 1.0, L)
        ray = np.zeros((L, M))
        ray[:, 0] = 1.0
        ray[:, 2] = 0.1
        mom = r[:, None] * ray + (1.0 - r[:, None]) * hg_moments(g, M)
        T = np.linspace(180.0, 300.0, L + 1)
        lo, hi = _bands(W)
        return Columns("c5_mixed_fwd_adj_N16", W, L, N, tau, ssa, mom, T, lo, hi,
                       surface_albedo=0.15, solar_flux=1.0, solar_mu=0.3)
    raise ValueError(f"unknown configuration {k}")


def planck_inputs_host(cols: Columns, planck_fn):
    """Planck levels, surface and top emission with ADConfig thermal semantics.

    ``planck_fn(lo, hi, T)`` is a scalar band-Planck function; used for small
    column sets on the host.  Returns (B [W][L+1], B_surface [W], B_top [W]).
    """
    W, L = cols.W, cols.L
    T = np.broadcast_to(cols.temperature, (W, L + 1))
    B = np.empty((W, L + 1))
    for w in range(W):
        for l in range(L + 1):
            B[w, l] = planck_fn(cols.wn_lo[w], cols.wn_hi[w], T[w, l])
    if cols.surface_temperature >= 0.0:
        Bs = np.array([planck_fn(cols.wn_lo[w], cols.wn_hi[w], cols.surface_temperature)
                       for w in range(W)])
    else:
        Bs = B[:, L].copy()
    Bt = B[:, 0].copy()
    return B, Bs, Bt


def batch_config(cols: Columns, **overrides):
    from .solver import BatchConfig
    kw = dict(num_quadrature=cols.N, use_delta_m=cols.use_delta_m,
              use_diffusion_lower_bc=cols.use_diffusion_lower_bc,
              surface_albedo=cols.surface_albedo, solar_flux=cols.solar_flux,
              solar_mu=cols.solar_mu)
    kw.update(overrides)
    return BatchConfig(**kw)


def device_inputs(cols: Columns, device="cuda"):
    """Device tensors for ``solve_batch``; Planck levels come from the device kernel."""
    import torch

    from .solver import planck_levels

    f64 = dict(dtype=torch.float64, device=device)
    d = dict(delta_tau=torch.as_tensor(cols.delta_tau, **f64).contiguous(),
             single_scat_albedo=torch.as_tensor(cols.single_scat_albedo, **f64).contiguous(),
             phase_moments=torch.as_tensor(cols.phase_moments, **f64).contiguous(),
             planck_levels=None, surface_emission=None, top_emission=None)
    if cols.thermal:
        lo = torch.as_tensor(cols.wn_lo, **f64)
        hi = torch.as_tensor(cols.wn_hi, **f64)
        T = torch.as_tensor(cols.temperature, **f64).contiguous()
        B = planck_levels(lo, hi, T)
        if cols.surface_temperature >= 0.0:
            Ts = torch.full((1,), cols.surface_temperature, **f64)
            Bs = planck_levels(lo, hi, Ts)[:, 0].contiguous()
        else:
            Bs = B[:, cols.L].contiguous()
        d.update(planck_levels=B, surface_emission=Bs, top_emission=B[:, 0].contiguous())
    return d
