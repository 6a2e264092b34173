"""The reduced-precision forms against the FP64 form and the FP64 oracle.

Bar (the accuracy the reduced-precision forms are held to): fluxes and mean
intensity within 1e-6 of FP64, per column relative to that output's largest
value in the column; gradients within 1e-4 of the FP64 reverse pass, per
column norm-wise (||g - g_f64|| / ||g_f64|| over the column's entries), and
for arrays summed over the columns (shared moments, surface albedo) as one
relative norm.

* "mixed_reverse": the FP64 forward (outputs identical to the plain form, bit
  for bit) with the reverse pass's doubling recomputation and its reverse in
  FP32 on the matrix cores.  Layers the FP32 reverse does not hold to the bar
  run the FP64 reverse kernels: layers with an unstable doubling start (omega
  < 0.01 with tau0 past the smallest quadrature cosine) and their scattering
  neighbours, and thick, strongly scattering layers (omega >= 0.8, tau >= 4).
  Held to the bar on every column.
* "mixed": the FP32 doubling in the forward as well (S = I - T form, FP64
  start, FP64-refined solves, FP64 adding and fluxes; the same layers in
  FP64).  On the 65536 x 100, N = 16 workload it leaves 23 % of the columns
  beyond 1e-6 in the fluxes (largest 3.6e-4, median 4.8e-7), so its tests at
  the bar are expected to fail there and on config 2 (strict xfail, with the
  measured figures printed) and pass on the N = 8 workloads (configs 1 and 4,
  and config 2's gradients); a regression bound at its measured ceiling
  guards it.  Its gradients of
  unstable-start layers (which both forms compute in FP64, but from interface
  vectors that differ by the FP32 rounding of the other layers' records) are
  reported apart from the others.  At quadrature order 32 (forward only) it
  misses the bar on most columns of config 3 and is held to a regression
  bound; layers it routes to FP64 come out bit for bit.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from addingdoublingrt_amd import adjoint_batch, solve_batch, workloads  # noqa: E402
from test_gpu_adjoint import unstable_layers  # noqa: E402

pytestmark = pytest.mark.gpu

WORKLOADS = [(1, 48), (2, 16), (4, 48), (5, 24), (5, 4096)]
FLUX_BAR = 1e-6
GRAD_BAR = 1e-4
OUTS = ("flux_up", "flux_down", "mean_intensity", "flux_direct")


def _inputs(k, W):
    cols = workloads.make(k, W=W)
    d = workloads.device_inputs(cols, "cuda")
    return cols, d, workloads.batch_config(cols)


def _run(cols, d, cfg, prec, seed):
    W, L = cols.W, cols.L
    g = torch.Generator(device="cuda").manual_seed(seed)
    cot = {n: torch.rand((W, L + 1), dtype=torch.float64, device="cuda", generator=g)
           for n in ("flux_up", "flux_down", "mean_intensity")}
    args = (d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg)
    kw = dict(planck_levels=d["planck_levels"], surface_emission=d["surface_emission"],
              top_emission=d["top_emission"])
    out = solve_batch(*args, **kw, adjoint_workspace=True, precision=prec)
    gen = out.pop("_generation")
    gr = adjoint_batch(*args, out, cot, **kw, forward_generation=gen, precision=prec)
    torch.cuda.synchronize()
    return ({n: v.cpu().numpy() for n, v in out.items()},
            {n: v.cpu().numpy() for n, v in gr.items() if v is not None})


def flux_errors(out, ref):
    """per column: the largest |x - x_ref| over interfaces and outputs,
    relative to that output's largest |x_ref| in the column"""
    e = 0.0
    for n in OUTS:
        scale = np.maximum(np.abs(ref[n]).max(axis=1), 1e-300)
        e = np.maximum(e, np.abs(out[n] - ref[n]).max(axis=1) / scale)
    return e


def grad_errors(g, ref, W, mask=None):
    """{name: per-column norm-wise relative differences} (one value for arrays
    summed over the columns); mask [W][L]: entries left out (delta_tau,
    single_scat_albedo, per-column moments)."""
    res = {}
    for n, b in ref.items():
        a = g[n]
        if n == "surface_albedo":
            res[n] = np.array([abs(a.sum() - b.sum()) / max(abs(b.sum()), 1e-300)])
            continue
        if not (n == "phase_moments" and b.ndim == 2):  # [W][...] per column
            a, b = a.reshape(W, -1).copy(), b.reshape(W, -1).copy()
            if mask is not None and n in ("delta_tau", "single_scat_albedo", "phase_moments"):
                m = mask if n != "phase_moments" else np.repeat(mask, ref[n].shape[-1], axis=1)
                a[m], b[m] = 0.0, 0.0
            nb = np.linalg.norm(b, axis=1)
            res[n] = np.where(nb > 0, np.linalg.norm(a - b, axis=1) / np.maximum(nb, 1e-300), 0.0)
        else:  # shared moments: one array
            res[n] = np.array([np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)])
    return res


@pytest.mark.parametrize("k,W", WORKLOADS)
def test_mixed_reverse_forward_is_the_f64_forward(k, W):
    cols, d, cfg = _inputs(k, W)
    o64, _ = _run(cols, d, cfg, "f64", 11 + k)
    omr, _ = _run(cols, d, cfg, "mixed_reverse", 11 + k)
    for n in o64:
        assert np.array_equal(o64[n], omr[n]), n


@pytest.mark.parametrize("k,W", WORKLOADS)
def test_mixed_reverse_gradients_within_bar(k, W):
    cols, d, cfg = _inputs(k, W)
    _, g64 = _run(cols, d, cfg, "f64", 11 + k)
    _, gmr = _run(cols, d, cfg, "mixed_reverse", 11 + k)
    for n, e in grad_errors(gmr, g64, W).items():
        print(f"mixed_reverse config {k} W={W} d/d{n}: per column max {e.max():.2e}, "
              f"median {np.median(e):.2e}")
        assert e.max() <= GRAD_BAR, (n, float(e.max()), int(e.argmax()))


@pytest.mark.parametrize("k,W", WORKLOADS)
def test_mixed_forward_regression(k, W):
    """The FP32-forward form at its measured ceiling (largest 3.6e-4 over all
    65536 columns of the headline workload, in a column with two
    unstable-start layers): a regression guard, not the bar."""
    cols, d, cfg = _inputs(k, W)
    out = solve_batch(d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg,
                      d["planck_levels"], d["surface_emission"], d["top_emission"],
                      precision="mixed", adjoint_workspace=True)
    out.pop("_generation")
    ref = solve_batch(d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg,
                      d["planck_levels"], d["surface_emission"], d["top_emission"])
    torch.cuda.synchronize()
    e = flux_errors({n: out[n].cpu().numpy() for n in OUTS}, {n: ref[n].cpu().numpy() for n in OUTS})
    print(f"mixed config {k} W={W}: fluxes per column max {e.max():.2e}, median "
          f"{np.median(e):.2e}, within 1e-6: {np.mean(e <= FLUX_BAR):.3f}")
    assert e.max() <= 4e-4, (float(e.max()), int(e.argmax()))


def _at_bar_cases(misses, reason):
    """WORKLOADS with the measured misses of the bar marked xfail (strict: a
    case that starts to pass is reported), the others held to it"""
    return [pytest.param(k, W, marks=pytest.mark.xfail(reason=reason, strict=True))
            if (k, W) in misses else (k, W) for k, W in WORKLOADS]


@pytest.mark.parametrize("k,W", _at_bar_cases(
    {(2, 16), (5, 24), (5, 4096)},
    "the FP32-forward form misses the 1e-6 flux bar on N = 16 workloads with thick or "
    "unstable-start layers (a quarter of the headline workload's columns, largest 3.6e-4)"))
def test_mixed_forward_vs_oracle_at_bar(k, W):
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
    e = flux_errors({n: out[n].cpu().numpy() for n in OUTS}, ref)
    print(f"mixed vs FP64 oracle config {k} W={W}: per column max {e.max():.2e}, "
          f"within 1e-6: {np.mean(e <= FLUX_BAR):.3f}")
    assert e.max() <= FLUX_BAR, (float(e.max()), int(e.argmax()))


@pytest.mark.parametrize("k,W", _at_bar_cases(
    {(5, 24), (5, 4096)},
    "the FP32-forward form's gradients miss 1e-4 on the headline workload, where its "
    "interface vectors differ from FP64 (unstable-start layers, thick layers)"))
def test_mixed_gradients_at_bar(k, W):
    cols, d, cfg = _inputs(k, W)
    _, g64 = _run(cols, d, cfg, "f64", 11 + k)
    _, gmx = _run(cols, d, cfg, "mixed", 11 + k)
    unstable = unstable_layers(cols.delta_tau, cols.single_scat_albedo, cols.phase_moments, cfg)
    worst = 0.0
    # reported apart: the entries of unstable-start layers, then all others
    for label, mask in (("without unstable-start layers", unstable), ("all entries", None)):
        for n, e in grad_errors(gmx, g64, cols.W, mask).items():
            print(f"mixed config {k} W={W} d/d{n} ({label}): per column max {e.max():.2e}")
            if mask is not None:
                worst = max(worst, float(e.max()))
    if unstable.any():
        a, b = gmx["delta_tau"][unstable], g64["delta_tau"][unstable]
        print(f"mixed config {k} W={W}: {int(unstable.sum())} unstable-start layers, their "
              f"d/ddelta_tau relative differences up to "
              f"{np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)):.2e}")
    assert worst <= GRAD_BAR, worst


# quadrature order 32 (workload config 3: 1 - omega down to 1e-6, 64 moments):
# the mixed forward alone (the reduced-precision reverse forms run up to order 16).  Measured on
# 8192 columns: per column largest 1.6e-4, median 7.7e-6, 5.5 % within 1e-6 --
# the floor that rounding the thin-layer start to FP32 already sets
# (tests/test_mixed32_ceiling.py) -- and no faster than FP64.
N32_COLUMNS = 512


def _mixed32_errors(W):
    cols, d, cfg = _inputs(3, W)
    args = (d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg,
            d["planck_levels"], d["surface_emission"], d["top_emission"])
    out = solve_batch(*args, precision="mixed")
    ref = solve_batch(*args)
    torch.cuda.synchronize()
    return cols, d, cfg, out, flux_errors({n: out[n].cpu().numpy() for n in OUTS},
                                          {n: ref[n].cpu().numpy() for n in OUTS})


def test_mixed32_forward_regression():
    """N = 32 mixed forward at its measured ceiling: finite everywhere, every
    column within 5e-4 of FP64 and the median within 3e-5 (a regression guard,
    not the bar)."""
    _, _, _, _, e = _mixed32_errors(N32_COLUMNS)
    print(f"mixed N=32 config 3 W={N32_COLUMNS}: per column max {e.max():.2e}, median "
          f"{np.median(e):.2e}, within 1e-6: {np.mean(e <= FLUX_BAR):.3f}")
    assert np.isfinite(e).all()
    assert e.max() <= 5e-4, (float(e.max()), int(e.argmax()))
    assert np.median(e) <= 3e-5, float(np.median(e))


def test_mixed32_routed_layers_are_fp64():
    """Columns whose every scattering layer the prep kernel routes to FP64
    (thick, strongly scattering: omega >= 0.8, tau >= 4) come out of the mixed
    form identical to the FP64 form, bit for bit."""
    cols, d, cfg = _inputs(3, 8)
    tau = d["delta_tau"].clone()
    tau.clamp_(min=4.5)  # every layer thick; config 3's omega >= 1 - 1e-1
    args = (tau, d["single_scat_albedo"], d["phase_moments"], cfg,
            d["planck_levels"], d["surface_emission"], d["top_emission"])
    assert float(d["single_scat_albedo"].min()) >= 0.8
    out = solve_batch(*args, precision="mixed")
    ref = solve_batch(*args)
    torch.cuda.synchronize()
    for n in OUTS:
        assert torch.equal(out[n], ref[n]), n


@pytest.mark.xfail(reason="FP32 doubling at N = 32 misses the 1e-6 flux bar on most columns of "
                          "config 3 (rounding the thin-layer start to FP32 alone moves the "
                          "fluxes by 3e-7 - 3e-6)", strict=False)
def test_mixed32_forward_vs_oracle_at_bar():
    cols, d, cfg, out, _ = _mixed32_errors(64)
    h = {n: (None if v is None else v.cpu().numpy()) for n, v in d.items()}
    ref = oracle.solve_batch(
        cols.delta_tau, cols.single_scat_albedo, cols.phase_moments, cols.N,
        planck_levels=h["planck_levels"],
        surface_emission=0.0 if h["surface_emission"] is None else h["surface_emission"],
        top_emission=0.0 if h["top_emission"] is None else h["top_emission"],
        surface_albedo=cfg.surface_albedo, solar_flux=cfg.solar_flux, solar_mu=cfg.solar_mu,
        use_delta_m=cfg.use_delta_m, use_diffusion_lower_bc=cfg.use_diffusion_lower_bc)
    e = flux_errors({n: out[n].cpu().numpy() for n in OUTS}, ref)
    print(f"mixed N=32 vs FP64 oracle config 3: per column max {e.max():.2e}, "
          f"within 1e-6: {np.mean(e <= FLUX_BAR):.3f}")
    assert e.max() <= FLUX_BAR, (float(e.max()), int(e.argmax()))
