"""Every quadrature order from 2 to 32, through the C-ABI.

The library runs orders 2, 4, 8, 16 and 32 on their own layouts and every
other order padded to the next of them with zero-weight nodes (mu = 1,
w = 0), the reverse pass at 16 (orders up to 16) or 32.  Per order, on a small
batch with the options rotated over the orders (solar beam with a surface,
delta-M with per-column moments, diffusion lower boundary, bottom-first
indexing, cold black surface):

* the forward against the FP64 oracle and the extended-precision oracle, in
  the accuracy class of tests/test_gpu_parity.py (1e-10 of the FP64 oracle on
  well-conditioned columns);
* the reverse pass's gradients of the thermal sources (Planck levels, surface
  and top emission) against the contraction of the cotangents with the source
  Jacobians (planck_source_vjp), a separate route through the adding: the
  outputs are affine in the sources, so both are exact up to rounding (1e-10);
* the reverse pass's gradients of tau, omega and the moments against central
  differences of the extended-precision oracle (the bar of
  tests/test_gpu_adjoint.py), on a smaller batch;
* at orders 2..16, the mixed_reverse form: outputs bit-identical to FP64,
  gradients within 1e-4 of the FP64 reverse pass per column.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from addingdoublingrt_amd import BatchConfig  # noqa: E402
from addingdoublingrt_amd.solver import planck_source_vjp  # noqa: E402
from test_gpu_adjoint import gpu_grads, random_cot, t  # noqa: E402
from test_gpu_parity import run_both  # noqa: E402

pytestmark = pytest.mark.gpu

OPTIONS = [
    dict(surface_albedo=0.2, solar_flux=1.0, solar_mu=0.6),
    dict(use_delta_m=True, surface_albedo=0.1, solar_flux=0.5, solar_mu=0.4),
    dict(use_diffusion_lower_bc=True, solar_flux=0.8, solar_mu=0.9),
    dict(index_from_bottom=True, surface_albedo=0.3, solar_flux=1.0, solar_mu=0.8),
    dict(surface_emission=0.0, solar_flux=1.0, solar_mu=0.5),
]


def order_batch(n, W=5, L=7, g_range=(0.2, 0.85)):
    rng = np.random.default_rng(700 + n)
    tau = 10.0 ** rng.uniform(-2.0, 1.0, (W, L))
    frac = np.log2(tau) % 1.0
    tau = np.where((frac < 1e-3) | (frac > 1 - 1e-3), tau * 1.01, tau)
    ssa = rng.uniform(0.1, 0.98, (W, L))
    g = rng.uniform(*g_range, (W, L))
    mom = g[..., None] ** np.arange(2 * n + 3)
    B = rng.uniform(0.5, 2.0, (W, L + 1))
    return tau, ssa, mom, B


@pytest.mark.parametrize("n", list(range(2, 33)))
def test_every_order(n):
    opt = OPTIONS[n % len(OPTIONS)]
    tau, ssa, mom, B = order_batch(n)
    W, L = tau.shape
    cfg = BatchConfig(num_quadrature=n, **opt)
    cold = opt.get("surface_emission") == 0.0
    Bs = None if cold or cfg.use_diffusion_lower_bc else B[:, -1] * 1.05
    Bt = B[:, 0] * 0.9
    run_both(tau, ssa, mom, cfg, B, Bs, Bt, label=f"order {n} {sorted(opt)}")
    cot = random_cot(W, L, n, ("flux_up", "flux_down", "mean_intensity", "flux_divergence"))
    g, _ = gpu_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt)
    vB, vS, vT = planck_source_vjp(t(tau), t(ssa), t(mom), cfg,
                                   {k: t(v) for k, v in cot.items()})
    vB = vB.cpu().numpy()
    np.testing.assert_allclose(g["planck_levels"], vB, rtol=1e-10, atol=1e-12 * np.abs(vB).max(),
                               err_msg=f"order {n}: d/dplanck_levels")
    np.testing.assert_allclose(g["top_emission"], vT.cpu().numpy(), rtol=1e-10, atol=1e-12,
                               err_msg=f"order {n}: d/dtop_emission")
    if Bs is not None:
        np.testing.assert_allclose(g["surface_emission"], vS.cpu().numpy(), rtol=1e-10,
                                   atol=1e-12, err_msg=f"order {n}: d/dsurface_emission")


@pytest.mark.parametrize("n", list(range(2, 33)))
def test_every_order_reverse_vs_differences(n):
    """The reverse pass at every order against central differences of the
    extended-precision oracle (the protocol and bar of tests/test_gpu_adjoint.py)
    for tau, omega and the per-column moments, on a small batch."""
    from test_gpu_adjoint import NAMES, assert_close, fd_grads, unstable_layers
    opt = OPTIONS[(n + 2) % len(OPTIONS)]
    # delta-M: the fraction f = g^(2n) stays above 1e-4, far from the 1e-12
    # switch of the reference's delta-M (a difference step must not cross it)
    g_range = (0.87, 0.93) if opt.get("use_delta_m") else (0.2, 0.85)
    tau, ssa, mom, B = order_batch(n, W=2, L=4, g_range=g_range)
    W, L = tau.shape
    cfg = BatchConfig(num_quadrature=n, **opt)
    cold = opt.get("surface_emission") == 0.0
    Bs = None if cold or cfg.use_diffusion_lower_bc else B[:, -1] * 1.05
    Bt = B[:, 0] * 0.9
    cot = random_cot(W, L, 100 + n, ("flux_up", "flux_down", "mean_intensity", "flux_direct"))
    g, _ = gpu_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt)
    fd, nz = fd_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt, which=("tau", "ssa", "mom"))
    loose = unstable_layers(tau, ssa, mom, cfg)
    for k, v in fd.items():
        assert_close(g[NAMES[k]], v, f"order {n} d/d{NAMES[k]}", loose, skip_loose=(k == "mom"),
                     noise=nz[k])


@pytest.mark.parametrize("n", list(range(2, 17)))
def test_every_order_mixed_reverse(n):
    """The mixed_reverse form (FP32 reverse doubling) at every order it takes:
    outputs bit-identical to the FP64 form's, gradients within the 1e-4 bar of
    the FP64 reverse pass, per column and norm-wise."""
    from addingdoublingrt_amd import adjoint_batch, solve_batch
    from test_gpu_mixed import grad_errors
    opt = OPTIONS[n % len(OPTIONS)]
    tau, ssa, mom, B = order_batch(n, W=16, L=12)
    W, L = tau.shape
    cfg = BatchConfig(num_quadrature=n, **opt)
    args = (t(tau), t(ssa), t(mom), cfg)
    kw = dict(planck_levels=t(B), top_emission=t(B[:, 0] * 0.9))
    cot = {k: t(v) for k, v in random_cot(W, L, 200 + n, ("flux_up", "flux_down",
                                                          "mean_intensity")).items()}
    res = {}
    for prec in ("f64", "mixed_reverse"):
        out = solve_batch(*args, **kw, adjoint_workspace=True, precision=prec)
        gen = out.pop("_generation")
        g = adjoint_batch(*args, out, cot, **kw, forward_generation=gen, precision=prec)
        torch.cuda.synchronize()
        res[prec] = ({k: v.cpu().numpy() for k, v in out.items()},
                     {k: v.cpu().numpy() for k, v in g.items() if v is not None})
    for k, v in res["f64"][0].items():
        assert np.array_equal(v, res["mixed_reverse"][0][k]), f"order {n}: {k}"
    errs = grad_errors(res["mixed_reverse"][1], res["f64"][1], W)
    worst = {k: float(v.max()) for k, v in errs.items()}
    print(f"order {n} mixed_reverse gradients, worst column: {worst}")
    for k, v in worst.items():
        assert v <= 1e-4, (n, k, v)
