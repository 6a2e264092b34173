"""Mixed-precision form (precision="mixed", N <= 16 on a reverse-pass
workspace) against the FP64 oracle and the FP64 GPU reverse pass.

The mixed form runs the doubling in FP32 on matrix cores in S = I - T form,
with the (I - R^2)^-1 solve refined by one FP64-residual step, and keeps the
thin-layer start, the source vectors, the adding and the interface fluxes in
FP64; layers whose reference start is unstable (omega < 0.01 with tau0 past
the smallest quadrature cosine) are recomputed by the FP64 kernel.  Its reverse
pass runs the recomputed doubling and its reverse in FP32 with the same FP64
parts.

Accuracy measured on these workloads (relative to |oracle| with a floor of
1e-12 of the column maximum):
  flux_up / flux_down   max 5.7e-6, median 5e-7 .. 2e-6
  mean_intensity        max 7.2e-6 here; up to 3.0e-5 measured on other draws of
                        the 65536 x 100 workload, in columns with unstable-start layers
Gradients against the FP64 reverse pass, norm-wise per column
(||g_mixed - g_f64|| / ||g_f64|| over each column's entries), leaving out the
entries of unstable-start layers (below):
  Planck levels, surface and top emission   <= 2.5e-6
  delta_tau                                  <= 4.9e-5
  phase moments                              <= 9.8e-5
  surface albedo                             <= 1.3e-4
  single_scat_albedo                         <= 1.4e-3 (N = 16, omega ~ U(0, 1), tau up to 10)
The bounds below are those figures with a margin of 2 to 2.5; they are the
accuracy this form delivers (fluxes short of the 1e-6 it was meant to reach).

On 4096 columns of the 65536 x 100 workload, against the FP64 form and
relative to each column's largest value (test_mixed_unstable_columns):
fluxes at most 2.2e-5 and mean intensity 6.8e-5, in columns with
unstable-start layers.  Before the layers redone in FP64 kept FP64 records
the mean intensity reached 2.6e-4 there (9.3e-3 over all 65536 columns).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from addingdoublingrt_amd import adjoint_batch, solve_batch, workloads  # noqa: E402
from test_gpu_adjoint import unstable_layers  # noqa: E402

pytestmark = pytest.mark.gpu

WORKLOADS = [(1, 48), (2, 16), (4, 48), (5, 24)]
FLUX_BOUND = 1.2e-5
MEAN_INTENSITY_BOUND = 6e-5
GRAD_BOUND = {"planck_levels": 6e-6, "surface_emission": 6e-6, "top_emission": 6e-6,
              "delta_tau": 1.2e-4, "phase_moments": 2.5e-4, "single_scat_albedo": 3.5e-3,
              "surface_albedo": 3e-4}


def _inputs(k, W):
    cols = workloads.make(k, W=W)
    d = workloads.device_inputs(cols, "cuda")
    return cols, d, workloads.batch_config(cols)


@pytest.mark.parametrize("k,W", WORKLOADS)
def test_mixed_forward_vs_oracle(k, W):
    cols, d, cfg = _inputs(k, W)
    out = solve_batch(d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg,
                      d["planck_levels"], d["surface_emission"], d["top_emission"],
                      precision="mixed", adjoint_workspace=True)
    out.pop("_generation")
    torch.cuda.synchronize()
    h = {n: (None if v is None else v.cpu().numpy()) for n, v in d.items()}
    ref = oracle.solve_batch(
        cols.delta_tau, cols.single_scat_albedo, cols.phase_moments, cols.N,
        planck_levels=h["planck_levels"],
        surface_emission=0.0 if h["surface_emission"] is None else h["surface_emission"],
        top_emission=0.0 if h["top_emission"] is None else h["top_emission"],
        surface_albedo=cfg.surface_albedo, solar_flux=cfg.solar_flux, solar_mu=cfg.solar_mu,
        use_delta_m=cfg.use_delta_m, use_diffusion_lower_bc=cfg.use_diffusion_lower_bc)
    for name, bound in (("flux_up", FLUX_BOUND), ("flux_down", FLUX_BOUND),
                        ("mean_intensity", MEAN_INTENSITY_BOUND), ("flux_direct", FLUX_BOUND)):
        g, r = out[name].cpu().numpy(), ref[name]
        scale = np.maximum(np.abs(r), 1e-12 * np.abs(r).max(axis=1, keepdims=True) + 1e-300)
        err = np.abs(g - r) / scale
        print(f"mixed config {k} {name}: max rel err {err.max():.2e}, median {np.median(err):.2e}")
        assert err.max() <= bound, (name, float(err.max()), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize("k,W", WORKLOADS)
def test_mixed_gradients_vs_f64(k, W):
    cols, d, cfg = _inputs(k, W)
    L = cols.L
    g = torch.Generator(device="cuda").manual_seed(11 + k)
    cot = {n: torch.rand((W, L + 1), dtype=torch.float64, device="cuda", generator=g)
           for n in ("flux_up", "flux_down", "mean_intensity")}
    args = (d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg)
    kw = dict(planck_levels=d["planck_levels"], surface_emission=d["surface_emission"],
              top_emission=d["top_emission"])
    res = {}
    for prec in ("f64", "mixed"):
        out = solve_batch(*args, **kw, adjoint_workspace=True, precision=prec)
        gen = out.pop("_generation")
        gr = adjoint_batch(*args, out, cot, **kw, forward_generation=gen, precision=prec)
        torch.cuda.synchronize()
        res[prec] = {n: v.cpu().numpy() for n, v in gr.items() if v is not None}
    unstable = unstable_layers(cols.delta_tau, cols.single_scat_albedo, cols.phase_moments, cfg)
    for name, bound in GRAD_BOUND.items():
        if name not in res["f64"]:
            continue
        a, b = res["mixed"][name], res["f64"][name]
        # entries of unstable-start layers carry derivatives that cancel terms
        # orders of magnitude larger (FP64 recovers about one digit of them, see
        # test_gpu_adjoint); both forms compute those layers with the FP64
        # kernels, so they are left out of the comparison
        if name in ("delta_tau", "single_scat_albedo"):
            a, b = np.where(unstable, 0.0, a), np.where(unstable, 0.0, b)
        elif name == "phase_moments":
            m = unstable[..., None] if b.ndim == 3 else unstable.any(axis=0)[:, None]
            a, b = np.where(m, 0.0, a), np.where(m, 0.0, b)
        if b.ndim and b.shape[0] == W:
            a, b = a.reshape(W, -1), b.reshape(W, -1)
        else:  # shared moments: one vector
            a, b = a.reshape(1, -1), b.reshape(1, -1)
        nb = np.linalg.norm(b, axis=1)
        live = nb > 0
        if not live.any():
            continue
        rel = np.linalg.norm(a - b, axis=1)[live] / nb[live]
        print(f"mixed config {k} d/d{name}: norm-wise per column max {rel.max():.2e}")
        assert rel.max() <= bound, (name, float(rel.max()))


def test_mixed_unstable_columns():
    """The mixed form against the FP64 form on 4096 columns of the 65536 x 100
    workload, about half of which hold an unstable-start layer."""
    cols, d, cfg = _inputs(5, 4096)
    args = (d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg,
            d["planck_levels"], d["surface_emission"], d["top_emission"])
    res = {}
    for prec in ("f64", "mixed"):
        out = solve_batch(*args, precision=prec, adjoint_workspace=True)
        out.pop("_generation")
        torch.cuda.synchronize()
        res[prec] = {n: v.cpu().numpy() for n, v in out.items()}
    for name, bound in (("flux_up", 5e-5), ("flux_down", 5e-5), ("mean_intensity", 1.5e-4)):
        a, b = res["mixed"][name], res["f64"][name]
        err = (np.abs(a - b) / np.abs(b).max(axis=1, keepdims=True)).max(axis=1)
        print(f"mixed unstable columns {name}: max {err.max():.2e}, median {np.median(err):.2e}")
        assert err.max() <= bound, (name, float(err.max()), int(err.argmax()))
