"""Hand-written reverse pass vs the CPU oracle, through the C-ABI.

Checks, each on the same seeded inputs:

* the exact forward-mode temperature Jacobian of the oracle (a restatement of
  the reference's computeTemperatureJacobian, src/solver.cpp:117-341):
  u^T J against the adjoint gradient chained with the device dB/dT, to
  1e-8 relative -- both sides are exact derivatives, no step error;
* central finite differences with a relative step of 1e-6 (the protocol of
  the reference's tests/test_temperature_jacobian.cpp:39-96) of the oracle
  run in x87 extended precision -- the same algorithm as the FP64 oracle, whose
  own rounding would otherwise reach 1e-4..1e-3 of small gradient components
  through the 1/step factor.  The oracle returns FP64 outputs, so the step is
  1e-6 * max(|x|, floor) with a floor (0.05 for tau and omega, 1 for moments
  and Planck values) that keeps one output ulp below 1e-4 of the gradient.
  Every differentiable input is checked, on the config-4 workload (thermal,
  Rayleigh, N = 8) and on option cases (solar beam, delta-M with per-column
  moments, bottom-first indexing, diffusion lower boundary, shared moments,
  divergence cotangents, spectrum scaling).  Bar: |g - g_fd| <= 1e-4 |g_fd|,
  with an absolute floor of 1e-6 max|g_fd| over the array for components that
  are negligible against the largest;
* layers where the reference's doubling start is unstable -- omega < 0.01
  starts at tau0 = tau / 2^(trunc(log2 tau) + 4) (src/doubling.hpp:20-26,
  107-112), and once tau0 exceeds the smallest quadrature cosine the thin-layer
  transmission 1 - tau0 / mu is negative and the doubling amplifies it (at
  N = 16 the layer's T[0][0] comes out near -1e4 instead of e^-370) -- are
  differentiated faithfully but their derivatives cancel terms many orders
  larger, so an FP64 reverse pass recovers their tau and omega gradients only
  to ~1e-1 and their moment gradients not at all; their tau / omega entries are
  compared at 2e-1 relative, their moment entries are not compared, and every
  other entry is held to the bar above;
* the forward on a reverse-pass workspace (quadrature padded to 16) within
  the accuracy class of the plain forward;
* the torch.autograd path (adrt.solve) against adjoint_batch.

Inputs avoid the points where the forward itself is not differentiable
(tau = 0, omega = 0 or 1, tau at a power of two -- the doubling count jumps --
and omega at the 0.01 / 0.1 thresholds of the initial-thickness exponent).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from addingdoublingrt_amd import BatchConfig, adjoint_batch, solve, solve_batch, workloads  # noqa: E402
from addingdoublingrt_amd.solver import planck_levels  # noqa: E402

pytestmark = pytest.mark.gpu
OUTS = ("flux_up", "flux_down", "mean_intensity", "flux_direct", "flux_divergence")
REL = 1e-4
FLOOR = 1e-6
STEP = 1e-6
LOOSE = 2e-1  # layers with an unstable doubling start (see above)


def t(a):
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                                                  device="cuda")


def oracle_kw(cfg, B, Bs, Bt):
    return dict(planck_levels=B, surface_emission=cfg.surface_emission if Bs is None else Bs,
                top_emission=cfg.top_emission if Bt is None else Bt,
                surface_albedo=cfg.surface_albedo, solar_flux=cfg.solar_flux,
                solar_mu=cfg.solar_mu, use_delta_m=cfg.use_delta_m,
                use_diffusion_lower_bc=cfg.use_diffusion_lower_bc,
                index_from_bottom=cfg.index_from_bottom, spectrum_scaling=cfg.spectrum_scaling)


def column_loss(res, cot):
    return sum(np.sum(cot[k] * res[k], axis=1) for k in cot)


def gpu_grads(tau, ssa, mom, cfg, cot, B=None, Bs=None, Bt=None):
    out = solve_batch(t(tau), t(ssa), t(mom), cfg, t(B), t(Bs), t(Bt), adjoint_workspace=True)
    gen = out.pop("_generation")
    g = adjoint_batch(t(tau), t(ssa), t(mom), cfg, out, {k: t(v) for k, v in cot.items()},
                      t(B), t(Bs), t(Bt), forward_generation=gen)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in g.items()}, {k: v.cpu().numpy() for k, v in out.items()}


def fd_grads(tau, ssa, mom, cfg, cot, B=None, Bs=None, Bt=None, which=None):
    """Central differences of the extended-precision oracle, step STEP * max(|x|, scale).

    Returns (gradients, resolution): the resolution of each component is four
    ulps of the FP64 outputs' loss over the step, below which a difference
    quotient carries no information."""
    N = cfg.num_quadrature
    base = dict(tau=tau, ssa=ssa, mom=mom, B=B, Bs=Bs, Bt=Bt)
    ulp4 = 4.0 * np.finfo(np.float64).eps

    def losses(**over):
        a = {**base, **over}
        r = oracle.solve_batch(a["tau"], a["ssa"], a["mom"], N, extended=True,
                               **oracle_kw(cfg, a["B"], a["Bs"], a["Bt"]))
        return column_loss(r, cot), sum(np.sum(np.abs(cot[k] * r[k]), axis=1) for k in cot)

    res, noise = {}, {}
    for name in which:
        x = base[name]
        g = np.zeros_like(x)
        nz = np.zeros_like(x)
        scale = {"tau": 0.05, "ssa": 0.05, "mom": 1.0, "B": 1.0, "Bs": 1.0, "Bt": 1.0}[name]
        if name == "mom" and x.ndim == 2:  # shared: total loss
            for idx in np.ndindex(*x.shape):
                h = STEP * max(abs(x[idx]), scale)
                xp, xm = x.copy(), x.copy()
                xp[idx] += h
                xm[idx] -= h
                (lp, ap), (lm, am) = losses(mom=xp), losses(mom=xm)
                g[idx] = (lp.sum() - lm.sum()) / (2 * h)
                nz[idx] = ulp4 * max(ap.sum(), am.sum()) / (2 * h)
        elif x.ndim == 1:  # [W]
            h = STEP * np.maximum(np.abs(x), scale)
            (lp, ap), (lm, am) = losses(**{name: x + h}), losses(**{name: x - h})
            g = (lp - lm) / (2 * h)
            nz = ulp4 * np.maximum(ap, am) / (2 * h)
        else:  # per-column arrays: perturb one inner index for all columns at once
            for idx in np.ndindex(*x.shape[1:]):
                sl = (slice(None),) + idx
                h = STEP * np.maximum(np.abs(x[sl]), scale)
                xp, xm = x.copy(), x.copy()
                xp[sl] += h
                xm[sl] -= h
                (lp, ap), (lm, am) = losses(**{name: xp}), losses(**{name: xm})
                g[sl] = (lp - lm) / (2 * h)
                nz[sl] = ulp4 * np.maximum(ap, am) / (2 * h)
        res[name] = g
        noise[name] = nz
    return res, noise


def albedo_fd(tau, ssa, mom, cfg, cot, B, Bs, Bt):
    """d loss / d surface albedo by central differences; returns (value, resolution)."""
    h = STEP * max(cfg.surface_albedo, 1e-3)
    N = cfg.num_quadrature
    tot, mag = [], []
    for A in (cfg.surface_albedo + h, cfg.surface_albedo - h):
        c2 = BatchConfig(**{**cfg.__dict__, "surface_albedo": A})
        r = oracle.solve_batch(tau, ssa, mom, N, extended=True, **oracle_kw(c2, B, Bs, Bt))
        tot.append(column_loss(r, cot).sum())
        mag.append(sum(np.sum(np.abs(cot[k] * r[k])) for k in cot))
    return (tot[0] - tot[1]) / (2 * h), 4.0 * np.finfo(np.float64).eps * max(mag) / (2 * h)


def unstable_layers(tau, ssa, mom, cfg):
    """[W][L] mask of layers whose doubling start tau0 exceeds the smallest mu at omega < 0.01."""
    from addingdoublingrt_amd.flops import doubling_counts, scaled_optics
    ts, os_ = scaled_optics(tau, ssa, mom, cfg.num_quadrature, cfg.use_delta_m)
    nn, scat = doubling_counts(ts, os_)
    mu_min = oracle.gauss_legendre(cfg.num_quadrature)[0][0]
    with np.errstate(divide="ignore", invalid="ignore"):
        return scat & (os_ < 0.01) & (ts / 2.0 ** nn > mu_min)


def assert_close(g, ref, label, loose=None, skip_loose=False, noise=None):
    floor = np.full(g.shape, FLOOR * max(np.abs(ref).max(), 1e-300))
    if noise is not None:
        floor = floor + noise
    err = np.abs(g - ref)
    rel = np.full(g.shape, REL)
    if loose is not None and loose.any():
        # an unstable layer's entries: LOOSE relative to the entry or to the
        # largest entry of that layer
        for idx in np.argwhere(loose):
            sl = tuple(idx)
            rel[sl] = LOOSE
            floor[sl] = np.inf if skip_loose else np.maximum(floor[sl],
                                                              LOOSE * np.abs(ref[sl]).max())
    bound = rel * np.abs(ref) + floor
    bad = np.argwhere(err > bound)
    if len(bad):
        i = tuple(bad[0])
        raise AssertionError(f"{label}: {len(bad)} of {g.size} off; first {i}: adjoint {g[i]!r} "
                             f"fd {ref[i]!r} (bound {bound[i]:.2e})")
    worst = np.max(err / bound)
    print(f"{label}: max |g - fd| / bound = {worst:.2e}")
    return worst


NAMES = {"tau": "delta_tau", "ssa": "single_scat_albedo", "mom": "phase_moments",
         "B": "planck_levels", "Bs": "surface_emission", "Bt": "top_emission"}


def smooth_batch(n, W=5, L=7, seed=11, M=None, per_column_moments=True):
    """Random columns away from the non-differentiable points of the forward."""
    rng = np.random.default_rng(seed)
    M = M or 2 * n
    tau = 10.0 ** rng.uniform(-2.5, 0.8, (W, L))
    frac = np.log2(tau) % 1.0
    tau = np.where((frac < 1e-3) | (frac > 1 - 1e-3), tau * 1.01, tau)
    ssa = rng.uniform(0.15, 0.95, (W, L))
    ssa[0, 1] = 0.05  # ipow0 = 10 branch
    ssa[1, 2] = 0.004  # ipow0 = 4 branch
    g = rng.uniform(0.2, 0.8, (W, L) if per_column_moments else (L,))
    mom = g[..., None] ** np.arange(M)
    B = rng.uniform(0.5, 2.0, (W, L + 1))
    return tau, ssa, mom, B


def random_cot(W, L, seed, keys=("flux_up", "flux_down", "mean_intensity")):
    rng = np.random.default_rng(seed)
    return {k: rng.uniform(-1, 1, (W, L + 1)) for k in keys}


# --------------------------------------------------------------------------


def test_adjoint_workspace_forward_matches_plain():
    from test_gpu_parity import run_both  # accuracy class vs the extended-precision oracle
    for n in (4, 8, 16):
        tau, ssa, mom, B = smooth_batch(n, W=9, L=8, seed=n)
        cfg = BatchConfig(num_quadrature=n, surface_albedo=0.3, solar_flux=2.0, solar_mu=0.7)
        run_both(tau, ssa, mom, cfg, B, B[:, -1].copy(), B[:, 0].copy(),
                 label=f"reverse-pass workspace N={n}", adjoint_workspace=True)


@pytest.mark.parametrize("n,delta_m", [(8, False), (4, True), (16, False), (2, False), (3, False),
                                        (6, True)])
def test_adjoint_matches_oracle_temperature_jacobian(n, delta_m):
    """u^T dF/dT from the reverse pass vs the oracle's forward-mode Jacobian."""
    rng = np.random.default_rng(7 + n)
    W, L = 4, 9
    tau, ssa, mom, _ = smooth_batch(n, W=W, L=L, seed=3 + n, M=2 * n + (3 if delta_m else 0))
    T = np.linspace(320.0, 200.0, L + 1)[None, :] + rng.normal(0, 5, (W, L + 1))
    Ts = rng.uniform(300, 320, W)
    lo = np.linspace(500, 1500, W)
    hi = lo + 40.0
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.1, use_delta_m=delta_m)
    B = planck_levels(t(lo), t(hi), t(T)).cpu().numpy()
    dBdT = planck_levels(t(lo), t(hi), t(T), derivative=True).cpu().numpy()
    Bs = np.array([oracle.planck(lo[w], hi[w], Ts[w]) for w in range(W)])
    dBsdT = np.array([oracle.planck_deriv(lo[w], hi[w], Ts[w]) for w in range(W)])
    cot = random_cot(W, L, 99, keys=("flux_up", "flux_down", "mean_intensity", "flux_divergence"))
    g, _ = gpu_grads(tau, ssa, mom, cfg, cot, B, Bs, B[:, 0].copy())
    worst = 0.0
    for w in range(W):
        r = oracle.solve(L, n, tau[w], ssa[w], mom[w], temperature=T[w], use_thermal_emission=True,
                         use_delta_m=delta_m, compute_temperature_jacobian=True,
                         surface_albedo=cfg.surface_albedo, surface_temperature=Ts[w],
                         wavenumber_low=lo[w], wavenumber_high=hi[w])
        ref = sum(cot[k][w] @ r[k + "_temperature_jac"] for k in
                  ("flux_up", "flux_down", "mean_intensity", "flux_divergence"))
        mine = g["planck_levels"][w] * dBdT[w]
        mine[0] += g["top_emission"][w] * dBdT[w, 0]
        mine = np.concatenate([mine, [g["surface_emission"][w] * dBsdT[w]]])
        scale = np.abs(ref).max()
        err = np.abs(mine - ref) / (np.abs(ref) + 1e-10 * scale)
        worst = max(worst, err.max())
        assert err.max() < 1e-8, (w, err.max(), mine, ref)
    print(f"N={n} delta_m={delta_m}: u^T J_T rel err {worst:.2e}")


def test_adjoint_vs_fd_config4():
    cols = workloads.make(4, W=4)
    d = workloads.device_inputs(cols, "cuda")
    h = {k: (None if v is None else v.cpu().numpy()) for k, v in d.items()}
    cfg = workloads.batch_config(cols)
    W, L = cols.W, cols.L
    rng = np.random.default_rng(4)
    cot = {"flux_up": np.zeros((W, L + 1))}
    lev0 = L if cfg.index_from_bottom else 0
    cot["flux_up"][:, lev0] = rng.uniform(0, 1, W)  # loss = sum_w u_w F_up(TOA)
    tau, ssa, mom = cols.delta_tau, cols.single_scat_albedo, cols.phase_moments
    B, Bs, Bt = h["planck_levels"], h["surface_emission"], h["top_emission"]
    g, _ = gpu_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt)
    fd, nz = fd_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt, which=("tau", "ssa", "B", "Bs", "Bt"))
    loose = unstable_layers(tau, ssa, mom, cfg)
    for k, v in fd.items():
        assert_close(g[NAMES[k]], v, f"config 4 d/d{NAMES[k]}",
                     loose if k in ("tau", "ssa") else None, noise=nz[k])
    gA, nA = albedo_fd(tau, ssa, mom, cfg, cot, B, Bs, Bt)
    got = g["surface_albedo"].sum()
    assert abs(got - gA) <= REL * abs(gA) + nA, (got, gA, nA)


CASES = {
    "solar_thermal": dict(n=8, cfg=dict(surface_albedo=0.3, solar_flux=2.0, solar_mu=0.6)),
    "delta_m": dict(n=4, M=11, cfg=dict(surface_albedo=0.2, solar_flux=1.0, solar_mu=0.5,
                                        use_delta_m=True)),
    "reversed": dict(n=8, cfg=dict(surface_albedo=0.25, solar_flux=1.0, solar_mu=0.8,
                                   index_from_bottom=True)),
    "diffusion": dict(n=8, cfg=dict(use_diffusion_lower_bc=True, solar_flux=0.5, solar_mu=0.9)),
    "shared_moments": dict(n=4, shared=True, cfg=dict(surface_albedo=0.15, solar_flux=1.0,
                                                      solar_mu=0.4)),
    "divergence_scaling": dict(n=4, div=True, cfg=dict(surface_albedo=0.1, solar_flux=1.0,
                                                       solar_mu=0.7, spectrum_scaling=0.6)),
    "n16": dict(n=16, L=4, cfg=dict(surface_albedo=0.2, solar_flux=1.0, solar_mu=0.5)),
    "no_surface": dict(n=4, cfg=dict(solar_flux=1.0, solar_mu=0.5), cold_surface=True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint_vs_fd_options(case):
    c = CASES[case]
    n = c["n"]
    L = c.get("L", 6)
    tau, ssa, mom, B = smooth_batch(n, W=4, L=L, seed=sum(map(ord, case)), M=c.get("M"),
                                    per_column_moments=not c.get("shared", False))
    cfg = BatchConfig(num_quadrature=n, **c["cfg"])
    Bs = None if c.get("cold_surface") else B[:, -1] * 1.1
    Bt = B[:, 0] * 0.9
    keys = ("flux_up", "flux_down", "mean_intensity", "flux_direct")
    if c.get("div"):
        keys = keys + ("flux_divergence",)
    cot = random_cot(tau.shape[0], L, 5, keys)
    if Bs is None:
        cfg = BatchConfig(**{**cfg.__dict__, "surface_emission": 0.0})
    g, _ = gpu_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt)
    which = ["tau", "ssa", "mom", "B", "Bt"] + ([] if Bs is None else ["Bs"])
    fd, nz = fd_grads(tau, ssa, mom, cfg, cot, B, Bs, Bt, which=which)
    loose = unstable_layers(tau, ssa, mom, cfg)
    for k, v in fd.items():
        lk = None
        if k in ("tau", "ssa") or (k == "mom" and mom.ndim == 3):
            lk = loose
        elif k == "mom":  # shared moments: a row is loose if any column's layer is
            lk = loose.any(axis=0)
        assert_close(g[NAMES[k]], v, f"{case} d/d{NAMES[k]}", lk, skip_loose=(k == "mom"),
                     noise=nz[k])
    if not cfg.use_diffusion_lower_bc and cfg.surface_albedo > 0.0:
        gA, nA = albedo_fd(tau, ssa, mom, cfg, cot, B, Bs, Bt)
        got = g["surface_albedo"].sum()
        assert abs(got - gA) <= REL * abs(gA) + nA, (got, gA, nA)


def test_autograd_matches_adjoint_batch():
    n = 8
    tau, ssa, mom, B = smooth_batch(n, W=6, L=5, seed=21)
    cfg = BatchConfig(num_quadrature=n, solar_flux=1.0, solar_mu=0.6)
    x = {k: t(v).requires_grad_(True) for k, v in dict(tau=tau, ssa=ssa, mom=mom, B=B).items()}
    A = torch.tensor(0.2, dtype=torch.float64, device="cuda", requires_grad=True)
    out = solve(x["tau"], x["ssa"], x["mom"], cfg, planck_levels=x["B"], surface_albedo=A)
    rng = np.random.default_rng(0)
    cot = {k: rng.uniform(-1, 1, (6, 6)) for k in ("flux_up", "flux_down", "mean_intensity")}
    loss = sum((out[k] * t(v)).sum() for k, v in cot.items())
    loss.backward()
    cfg2 = BatchConfig(**{**cfg.__dict__, "surface_albedo": 0.2})
    g, _ = gpu_grads(tau, ssa, mom, cfg2, cot, B)
    for k, name in (("tau", "delta_tau"), ("ssa", "single_scat_albedo"),
                    ("mom", "phase_moments"), ("B", "planck_levels")):
        np.testing.assert_allclose(x[k].grad.cpu().numpy(), g[name], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(A.grad.item(), g["surface_albedo"].sum(), rtol=1e-12)


def test_temperature_gradient_through_device_planck():
    """d loss / d T through planck_levels' autograd = chain with dB/dT."""
    W, L, n = 3, 4, 4
    tau, ssa, mom, _ = smooth_batch(n, W=W, L=L, seed=2)
    lo = t(np.array([600.0, 900.0, 1200.0]))
    hi = lo + 25.0
    T = t(np.linspace(280, 220, L + 1)).requires_grad_(True)
    Bl = planck_levels(lo, hi, T)
    out = solve(t(tau), t(ssa), t(mom), BatchConfig(num_quadrature=n), planck_levels=Bl)
    out["flux_up"][:, 0].sum().backward()
    # reference: the same chain done by hand with the oracle FD in T
    Th = T.detach().cpu().numpy()
    g = np.zeros(L + 1)
    for m in range(L + 1):
        h = 1e-6 * Th[m]
        vals = []
        for s in (1, -1):
            Tp = Th.copy()
            Tp[m] += s * h
            Bp = np.array([[oracle.planck(lo[w].item(), hi[w].item(), Tp[k]) for k in range(L + 1)]
                           for w in range(W)])
            r = oracle.solve_batch(tau, ssa, mom, n, planck_levels=Bp, extended=True)
            vals.append(r["flux_up"][:, 0].sum())
        g[m] = (vals[0] - vals[1]) / (2 * h)
    np.testing.assert_allclose(T.grad.cpu().numpy(), g, rtol=1e-4, atol=1e-6 * np.abs(g).max())


@pytest.mark.parametrize("n,diffusion", [(32, False), (32, True)])
def test_source_gradients_order_32(n, diffusion):
    """At order 32 the thermal-source gradients of solve() (the N = 32 reverse
    pass) against central differences of the GPU forward with a unit step,
    exact up to rounding since every output is affine in the sources, and
    against the source-Jacobian route (planck_source_vjp)."""
    W, L = 3, 5
    tau, ssa, mom, B = smooth_batch(n, W=W, L=L, seed=31)
    rng = np.random.default_rng(4)
    Bs, Bt = rng.uniform(0.5, 1.5, W), rng.uniform(0.1, 0.4, W)
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.0 if diffusion else 0.25,
                      use_diffusion_lower_bc=diffusion, solar_flux=1.0, solar_mu=0.6)
    xB, xS, xT = (t(v).requires_grad_(True) for v in (B, Bs, Bt))
    out = solve(t(tau), t(ssa), t(mom), cfg, planck_levels=xB, surface_emission=xS,
                top_emission=xT)
    names = ("flux_up", "flux_down", "mean_intensity", "flux_divergence")
    cot = {k: rng.uniform(-1, 1, (W, L + 1)) for k in names}
    (sum((out[k] * t(v)).sum() for k, v in cot.items())).backward()

    def loss(Bv, Sv, Tv):
        o = solve_batch(t(tau), t(ssa), t(mom), cfg, planck_levels=t(Bv),
                        surface_emission=t(Sv), top_emission=t(Tv))
        return sum(float((o[k].cpu().numpy() * v).sum()) for k, v in cot.items())

    def fd(which, shape_idx):
        args = [B.copy(), Bs.copy(), Bt.copy()]
        vals = []
        for s in (1.0, -1.0):
            a = [x.copy() for x in args]
            a[which][shape_idx] += s
            vals.append(loss(*a))
        return (vals[0] - vals[1]) / 2.0

    gB = xB.grad.cpu().numpy()
    for w in range(W):
        for m in range(L + 1):
            ref = fd(0, (w, m))
            assert abs(gB[w, m] - ref) <= 1e-10 * max(abs(ref), 1.0), (w, m, gB[w, m], ref)
        for which, g in ((1, xS.grad), (2, xT.grad)):
            ref = fd(which, w)
            got = g.cpu().numpy()[w]
            assert abs(got - ref) <= 1e-10 * max(abs(ref), 1.0), (which, w, got, ref)
    from addingdoublingrt_amd.solver import planck_source_vjp
    vB, vS, vT = planck_source_vjp(t(tau), t(ssa), t(mom), cfg, {k: t(v) for k, v in cot.items()})
    np.testing.assert_allclose(gB, vB.cpu().numpy(), rtol=1e-10, atol=1e-12 * np.abs(gB).max())
    np.testing.assert_allclose(xS.grad.cpu().numpy(), vS.cpu().numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(xT.grad.cpu().numpy(), vT.cpu().numpy(), rtol=1e-10, atol=1e-12)


def test_autograd_net_flux_with_direct_beam():
    """flux_direct = F0 mu0 exp(-tau_cum / mu0) with the delta-M-scaled depth is
    differentiable through solve(): a loss on the net flux F_up - F_down - F_dir
    against central differences of the extended-precision oracle, for tau,
    omega and (through delta-M) the per-column moments."""
    n, W, L = 4, 4, 6
    tau, ssa, mom, B = smooth_batch(n, W=W, L=L, seed=41, M=2 * n + 3)
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.2, solar_flux=1.5, solar_mu=0.55,
                      use_delta_m=True)
    rng = np.random.default_rng(8)
    u = rng.uniform(0.2, 1.0, (W, L + 1))
    x = {k: t(v).requires_grad_(True) for k, v in dict(tau=tau, ssa=ssa, mom=mom).items()}
    out = solve(x["tau"], x["ssa"], x["mom"], cfg, planck_levels=t(B))
    tu = t(u)
    ((out["flux_up"] - out["flux_down"] - out["flux_direct"]) * tu).sum().backward()
    cot = {"flux_up": u, "flux_down": -u, "flux_direct": -u}
    fd, nz = fd_grads(tau, ssa, mom, cfg, cot, B, which=("tau", "ssa", "mom"))
    loose = unstable_layers(tau, ssa, mom, cfg)
    for k, v in fd.items():
        assert_close(x[k].grad.cpu().numpy(), v, f"net flux d/d{NAMES[k]}", loose,
                     skip_loose=(k == "mom"), noise=nz[k])
    # the beam's part alone is the closed form's derivative: without it the
    # tau gradients would differ by -sum_l u_l dF_dir(l)/dtau
    x2 = {k: t(v).requires_grad_(True) for k, v in dict(tau=tau, ssa=ssa, mom=mom).items()}
    o2 = solve(x2["tau"], x2["ssa"], x2["mom"], cfg, planck_levels=t(B))
    (o2["flux_direct"] * tu).sum().backward()
    gd = x2["tau"].grad.cpu().numpy()
    assert np.abs(gd).max() > 1e-3 * np.abs(fd["tau"]).max()


def test_backward_after_workspace_rebuild():
    """Two differentiable solves of one column shape, the second larger so that
    the cached workspace is rebuilt in between: each backward uses the
    workspace that holds its own forward, and matches a separate run."""
    n, L = 8, 5
    tau, ssa, mom, B = smooth_batch(n, W=9, L=L, seed=77)
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.2, solar_flux=1.0, solar_mu=0.6)
    rng = np.random.default_rng(12)
    cot = {"flux_up": rng.uniform(-1, 1, (9, L + 1)), "mean_intensity": rng.uniform(-1, 1, (9, L + 1))}
    xs = {k: t(v[:2]).requires_grad_(True) for k, v in dict(tau=tau, ssa=ssa).items()}
    xl = {k: t(v).requires_grad_(True) for k, v in dict(tau=tau, ssa=ssa).items()}
    # a shape of its own, so that the small solve builds a workspace sized for 2 columns
    cfg_s = BatchConfig(**{**cfg.__dict__})
    o_small = solve(xs["tau"], xs["ssa"], t(mom[:2]), cfg_s, planck_levels=t(B[:2]))
    o_large = solve(xl["tau"], xl["ssa"], t(mom), cfg, planck_levels=t(B))
    loss = sum((o_small[k] * t(v[:2])).sum() + (o_large[k] * t(v)).sum() for k, v in cot.items())
    loss.backward()
    for sl, x in ((slice(0, 2), xs), (slice(0, 9), xl)):
        g, _ = gpu_grads(tau[sl], ssa[sl], mom[sl], cfg, {k: v[sl] for k, v in cot.items()},
                         B[sl])
        for k, name in (("tau", "delta_tau"), ("ssa", "single_scat_albedo")):
            np.testing.assert_allclose(x[k].grad.cpu().numpy(), g[name], rtol=1e-12,
                                       atol=1e-15 * np.abs(g[name]).max())
