"""Pair mode: two layers of quadrature order <= 8 per wave (adrt_pair16.hpp).

Reverse-pass workspaces run orders up to 8 padded to 16, and layer16p_kernel /
layer16p_adjoint_kernel put the real blocks of two layers of a column into one
block-diagonal 16 x 16 problem (layer A in rows and columns 0..7, layer B in
8..15; two matrix instructions per product).  Checked here, through the C-ABI:

* the forward on a reverse-pass workspace against the oracle, in the accuracy
  class of tests/test_gpu_parity.py, at N = 2, 3, 6, 8, with optical depths
  from 1e-4 to 50 side by side (doubling counts differing by up to 19, so one
  layer of a pair may be stored and frozen while the other still doubles), a
  non-scattering layer inside a column and an odd number of layers;
* gradients of tau, omega and the Planck levels against central differences
  of the extended-precision oracle (tests/test_gpu_adjoint.py's protocol and
  1e-4 bar) on such columns (scattering layers only: omega = 0 is a kink of
  the forward) at N = 3, 6 (delta-M) and 8;
* the reverse pass at N = 2, 3 and 6 against the oracle's exact temperature
  Jacobian (tests/test_gpu_adjoint.py, u^T J_T at 1e-8).

Pairs are taken within a column, its layers sorted by doubling count; the
last layer of an odd column and pairs holding a non-scattering layer run one
layer per wave.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from addingdoublingrt_amd import BatchConfig  # noqa: E402

pytestmark = pytest.mark.gpu


def spread_batch(n, W, L, seed, M=None):
    rng = np.random.default_rng(seed)
    M = M or 2 * n
    # alternate thin and thick layers: neighbouring tasks double 4..23 times
    lo = rng.uniform(-4.0, -2.0, (W, L))
    hi = rng.uniform(0.5, np.log10(50.0), (W, L))
    tau = 10.0 ** np.where((np.arange(L) % 2 == 0)[None, :], lo, hi)
    ssa = rng.uniform(0.15, 0.95, (W, L))
    g = rng.uniform(0.2, 0.8, (W, L))
    mom = g[..., None] ** np.arange(M)
    B = np.linspace(0.5, 2.0, L + 1)[None, :] * rng.uniform(0.8, 1.2, (W, 1))
    return tau, ssa, mom, B


@pytest.mark.parametrize("n", [2, 3, 6, 8])
def test_pair_forward_matches_oracle(n):
    from test_gpu_parity import run_both

    W, L = 5, 7  # odd columns: one layer of each runs alone
    tau, ssa, mom, B = spread_batch(n, W, L, seed=40 + n)
    ssa[1, 3] = 0.0  # purely absorbing layers: their pairs run one layer at a time
    ssa[3, 0] = 0.0
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.25, solar_flux=1.5, solar_mu=0.55)
    run_both(tau, ssa, mom, cfg, B, B[:, -1].copy(), B[:, 0].copy(),
             label=f"pair mode N={n}", adjoint_workspace=True)


@pytest.mark.parametrize("n,delta_m", [(3, False), (6, True), (8, False)])
def test_pair_gradients_vs_fd(n, delta_m):
    """Every input's gradient against central differences of the extended-precision
    oracle (tests/test_gpu_adjoint.py's protocol and bar), on columns whose
    sorted pairs double 4..23 times side by side (odd columns: one layer each runs
    alone)."""
    from test_gpu_adjoint import NAMES, assert_close, fd_grads, gpu_grads, random_cot

    W, L = 2, 7
    tau, ssa, mom, B = spread_batch(n, W, L, seed=90 + n, M=2 * n + (3 if delta_m else 0))
    frac = np.log2(tau) % 1.0  # away from the jumps of the doubling count
    tau = np.where((frac < 1e-3) | (frac > 1 - 1e-3), tau * 1.01, tau)
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.2, solar_flux=1.0, solar_mu=0.6,
                      use_delta_m=delta_m)
    cot = random_cot(W, L, 17 + n)
    Bs, Bt = B[:, -1].copy(), B[:, 0].copy()
    g, _ = gpu_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt)
    fd, nz = fd_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt, which=("tau", "ssa", "B"))
    for k, v in fd.items():
        assert_close(g[NAMES[k]], v, f"pair mode N={n} d/d{NAMES[k]}", noise=nz[k])
