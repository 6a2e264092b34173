"""Algorithmic work of one forward solve, counted from the actual inputs.

Per column (N = quadrature order, counted as dense N x N x N products):
  doubling step      (2 + 2/3 + 2 + 2 + 2 + 2) N^3   R^2, LU(A^T) + N-RHS solve, TG R,
                                                      TGR T, TG T  (src/doubling.hpp:149-209)
  general add        (2 + 2 + 4/3 + 4 + 8 + 4) N^3   two products, two LU + solves,
                                                      four products  (src/adding.hpp:69-110)
  interface solve    (4 + 4/3) N^3                   two products + two LUs
                                                      (src/solver.cpp:24-73)
O(N^2) terms (vectors, diagonal scalings, phase setup) are not counted, and the
count follows the reference algorithm whatever the kernels do internally.

Reverse pass (no counterpart in the reference): every forward product is
counted twice (dA = dC B^T, dB = A^T dC) and every right solve's reverse as a
solve plus a product with the transposed inverse (2 + 2) N^3; the doubling
steps recomputed from the checkpoints are not counted, so the reverse doubling
is (8 + 8 + 4) N^3 = 20 N^3 per step.  The adding reverse is one adding solve
of the transposed system, counted as the forward adding and interfaces.
"""
from __future__ import annotations

import numpy as np

DOUBLING = 2 + 2 / 3 + 2 + 2 + 2 + 2
DOUBLING_REVERSE = 2 * (2 + 2 + 2 + 2) + (2 + 2)
GENERAL_ADD = 2 + 2 + 4 / 3 + 4 + 8 + 4
INTERFACE = 4 + 4 / 3


def scaled_optics(tau, ssa, mom, N, use_delta_m):
    """Delta-M scaled (tau, omega) as src/solver.cpp:441-486 (vectorised)."""
    tau = np.asarray(tau, np.float64)
    ssa = np.asarray(ssa, np.float64)
    if not use_delta_m or mom.shape[-1] <= 2 * N:
        return tau, ssa
    f = np.broadcast_to(mom[..., 2 * N], tau.shape)
    active = (ssa > 0) & (tau > 0) & (f > 1e-12) & (f < 1 - 1e-12)
    omf = ssa * f
    t2 = np.where(active, (1.0 - omf) * tau, tau)
    o2 = np.where(active, ssa * (1.0 - f) / np.where(active, 1.0 - omf, 1.0), ssa)
    return t2, o2


def doubling_counts(tau_s, omega_s):
    """nn per layer (src/doubling.hpp:20-26, 107-109); 0 for non-scattering layers."""
    om = np.clip(omega_s, 0.0, 1.0)
    ip = np.where(om < 0.01, 4, np.where(om < 0.1, 10, 16))
    with np.errstate(divide="ignore", invalid="ignore"):
        nn = np.trunc(np.log(np.where(tau_s > 0, tau_s, 1.0)) / np.log(2.0)).astype(np.int64) + ip
    nn = np.maximum(nn, 1)
    scat = (tau_s > 0) & (omega_s > 0)
    return np.where(scat, nn, 0), scat


def forward_flops(tau, ssa, mom, N, use_delta_m=False, surface_albedo=0.0,
                  surface_emission=None, diffusion=False):
    """Returns dict(doubling=..., adding=..., interfaces=..., total=...) in FLOPs."""
    tau_s, om_s = scaled_optics(tau, ssa, mom, N, use_delta_m)
    nn, scat = doubling_counts(tau_s, om_s)
    W, L = tau_s.shape
    n3 = float(N) ** 3
    dbl = DOUBLING * n3 * float(nn.sum())
    if surface_emission is None:
        surface_emission = np.zeros(W)
    Bs = np.broadcast_to(np.asarray(surface_emission, np.float64), (W,))
    has_surface = (~diffusion) & ((surface_albedo > 0.0) | (Bs > 0.0))
    # composite scattering flag = OR of the layers it holds (src/adding.hpp:125,189),
    # so an add is "general" iff the new layer and the composite both scatter
    surf_col = (has_surface & (surface_albedo > 0.0))[:, None]
    flags = np.concatenate([scat, surf_col], axis=1)          # surface column (maybe unused)
    valid = np.concatenate([np.ones((W, L), bool), has_surface[:, None]], axis=1)
    f = flags & valid
    below = np.flip(np.maximum.accumulate(np.flip(f, axis=1), axis=1), axis=1)  # OR over k..end
    below_next = np.concatenate([below[:, 1:], np.zeros((W, 1), bool)], axis=1)
    n_general = int(np.sum(f[:, :-1] & below_next[:, :-1] & valid[:, :-1]))  # bottom-up
    above = np.maximum.accumulate(scat, axis=1)
    n_general += int(np.sum(scat[:, 1:] & above[:, :-1]))                       # top-down
    n_iface = int(np.sum(np.where(has_surface, L, L - 1)))
    add = GENERAL_ADD * n3 * n_general
    itf = INTERFACE * n3 * n_iface
    return dict(doubling=dbl, adding=add, interfaces=itf, total=dbl + add + itf,
                doubling_steps=int(nn.sum()), general_adds=n_general, interface_solves=n_iface,
                doubling_reverse=DOUBLING_REVERSE * n3 * float(nn.sum()))
