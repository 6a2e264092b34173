"""The reference's own test cases, run on the GPU batched path.

* Published-table cases (CDISORT, Van de Hulst, Sweigart, Garcia-Siewert,
  Ozisik-Shouman) that tests/test_ad_solver.cpp checks on the CPU solver and
  tests/test_cuda_solver.cu:194-1450 runs through the reference's batched CUDA
  solver, solved here by adrt_solve_batch: the published values at the
  reference's own tolerances, and the FP64 oracle at the parity bar (1e-10
  relative to the column maximum).
* Batch consistency (tests/test_cuda_solver.cu:1248-1296): 100 identical
  columns give identical results; here bit for bit (no reduction in the
  kernels depends on the batch).
* The raw-coefficient case of tests/test_cuda_solver.cu:1599-1760 (4 wavenumbers,
  5 layers, N = 4, hot profile 1500 -> 500 K): the raw entry against the batched
  solve on optics prepared on the host, and against the oracle.
* Closed forms and invariants of those test files (Beer's law with and without
  a Lambertian surface, the beam term of the mean intensity, the linear-in-tau
  pure-absorption source, flux conservation at omega = 1, delta-M leaving
  Rayleigh scattering unchanged, TOA downwelling pi B_top, zero divergence in
  an isothermal enclosure, the diffusion lower boundary's pi B + (2 pi / 3)
  dB/dtau and its independence of the albedo, the divergence integral against
  the net-flux change, thermal and solar, index reversal, workspace reuse,
  layer-split invariance) at N = 8, 16 and 32, and the convergence of orders
  6, 8, 12, 16 (6 and 12 padded with zero-weight nodes).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from addingdoublingrt_amd import BatchConfig, solve_batch, solve_batch_from_coefficients  # noqa: E402
from golden.reference_tables import CASES, THERMAL_CASES  # noqa: E402
from test_oracle_golden import _build  # noqa: E402

pytestmark = pytest.mark.gpu
OUTS = ("flux_up", "flux_down", "mean_intensity", "flux_direct")


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")


def _gpu_column(cfg):
    """One column of an oracle.solve configuration through the batched path."""
    L = cfg["num_layers"]
    M = max(len(m) for m in cfg["phase_function_moments"])
    mom = np.zeros((L, M))
    for l, m in enumerate(cfg["phase_function_moments"]):
        mom[l, :len(m)] = m
    bc = BatchConfig(num_quadrature=cfg["num_quadrature"], use_delta_m=cfg.get("use_delta_m", False),
                     surface_albedo=cfg.get("surface_albedo", 0.0),
                     surface_emission=cfg.get("surface_emission", 0.0),
                     top_emission=cfg.get("top_emission", 0.0),
                     solar_flux=cfg.get("solar_flux", 0.0), solar_mu=cfg.get("solar_mu", 1.0))
    B = cfg.get("planck_levels")
    out = solve_batch(_t(np.array(cfg["delta_tau"])[None]), _t(np.array(cfg["single_scat_albedo"])[None]),
                      _t(mom), bc, planck_levels=None if B is None else _t(np.array(B)[None]))
    torch.cuda.synchronize()
    return {k: v[0].cpu().numpy() for k, v in out.items()}


def _check(name, cfg, checks):
    g = _gpu_column(cfg)
    ref = oracle.solve(**cfg)
    for out, idx, expected, tol in checks:
        assert abs(g[out][idx] - expected) <= tol, (name, out, idx, g[out][idx], expected)
    for k in OUTS:
        scale = max(np.abs(ref[k]).max(), 1e-300)
        err = np.abs(g[k] - ref[k]).max() / scale
        assert err <= 1e-10, (name, k, err)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_published_tables_on_gpu(case):
    _check(case["name"], _build(case), case["checks"])


@pytest.mark.parametrize("case", THERMAL_CASES, ids=[c["name"] for c in THERMAL_CASES])
def test_published_thermal_tables_on_gpu(case):
    cfg = _build(case)
    cfg["planck_levels"] = [oracle.planck(*p) for p in case["planck"]]
    if case["top"] is not None:
        iso, p = case["top"]
        cfg["top_emission"] = iso + oracle.planck(*p)
    if case["surface"] is not None:
        cfg["surface_emission"] = oracle.planck(*case["surface"])
    _check(case["name"], cfg, case["checks"])


def test_batch_consistency():
    W, L, M = 100, 5, 3
    mom = np.zeros((L, M))
    mom[:, 0] = 1.0
    tau = np.full((W, L), 0.5)
    ssa = np.full((W, L), 0.8)
    B = np.tile(100.0 + 20.0 * np.arange(L + 1), (W, 1))
    out = solve_batch(_t(tau), _t(ssa), _t(mom), BatchConfig(num_quadrature=8, surface_albedo=0.1),
                      planck_levels=_t(B))
    torch.cuda.synchronize()
    for k in OUTS:
        v = out[k].cpu().numpy()
        assert np.array_equal(v, np.broadcast_to(v[0], v.shape)), k


def test_raw_entry_reference_case():
    nwav, nlay, N, nmom = 4, 5, 4, 8
    nu = np.array([500.0, 1000.0, 1500.0, 2000.0])
    z = np.array([0.0, 1e5, 3e5, 6e5, 1e6, 2e6])
    T = np.array([1500.0, 1300.0, 1100.0, 900.0, 700.0, 500.0])
    lev, w = np.meshgrid(np.arange(nlay + 1), np.arange(nwav), indexing="ij")
    a = 1e-6 * (1.0 + 0.5 * lev) * (1.0 + 0.3 * w)
    s = 0.3 * a
    mom = np.zeros((nlay, nmom))
    mom[:, 0] = 1.0
    cfg = BatchConfig(num_quadrature=N)
    raw = solve_batch_from_coefficients(_t(a), _t(s), _t(z), _t(T), _t(nu), _t(mom), cfg)
    r = oracle.raw_optics(a, s, z, T, nu)
    pre = solve_batch(_t(r["delta_tau"]), _t(r["single_scat_albedo"]), _t(mom), cfg,
                      planck_levels=_t(r["planck_levels"]),
                      surface_emission=_t(r["surface_emission"]),
                      top_emission=_t(r["top_emission"]))
    torch.cuda.synchronize()
    ref = oracle.solve_batch(r["delta_tau"], r["single_scat_albedo"], mom, N,
                             planck_levels=r["planck_levels"],
                             surface_emission=r["surface_emission"],
                             top_emission=r["top_emission"])
    for k in OUTS:
        g, p, o = raw[k].cpu().numpy(), pre[k].cpu().numpy(), ref[k]
        scale = np.abs(o).max(axis=1, keepdims=True)
        assert np.all(np.abs(g - p) <= 1e-13 * scale), k     # same solve, Planck to an ulp
        assert np.all(np.abs(g - o) <= 1e-10 * scale), k


# ---- closed forms and invariants the reference tests (tests/test_ad_solver.cpp,
# tests/test_cuda_solver.cu), on the GPU path at N = 8 (column slices), 16 and 32
# (matrix cores) ----

def _iso_cfg(nq, tau, ssa, **kw):
    n = len(tau)
    mom = [[1.0] + [0.0] * (2 * nq - 1)] * n
    return dict(num_layers=n, num_quadrature=nq, delta_tau=tau, single_scat_albedo=ssa,
                phase_function_moments=mom, **kw)


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_beers_law_on_gpu(nq):             # test_ad_solver.cpp:144-195
    g = _gpu_column(_iso_cfg(nq, [1.0], [0.0], solar_flux=200.0, solar_mu=0.5))
    assert abs(g["flux_direct"][0] - 100.0) < 1e-6
    assert abs(g["flux_direct"][1] - 100.0 * np.exp(-2.0)) < 1e-10
    assert np.all(np.abs(g["flux_up"]) < 1e-12) and np.all(np.abs(g["flux_down"]) < 1e-12)
    g = _gpu_column(_iso_cfg(nq, [1.0], [0.0], solar_flux=200.0, solar_mu=0.5,
                             surface_albedo=0.5))
    rfldir = 100.0 * np.exp(-2.0)
    assert abs(g["flux_up"][1] - 0.5 * rfldir) <= 1e-10 * rfldir
    assert 0.0 < g["flux_up"][0] < g["flux_up"][1]


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_mean_intensity_is_beam_term_on_gpu(nq):   # test_ad_solver.cpp:201-227
    g = _gpu_column(_iso_cfg(nq, [0.4, 0.6], [0.0, 0.0], solar_flux=100.0, solar_mu=0.6))
    tc = np.array([0.0, 0.4, 1.0])
    assert np.allclose(g["mean_intensity"], 100.0 * np.exp(-tc / 0.6) / (4 * np.pi),
                       rtol=1e-13, atol=0)


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_linear_source_pure_absorption_on_gpu(nq):  # test_ad_solver.cpp:1472-1510
    g = _gpu_column(_iso_cfg(nq, [1.0], [0.0], planck_levels=[1.0, 3.0]))
    mu, wt = oracle.gauss_legendre(nq)
    t = np.exp(-1.0 / mu)
    slope = mu * (1 - t) - 0.5 * (1 + t)
    Fu = np.sum(2 * np.pi * wt * mu * (2.0 * (1 - t) + 2.0 * slope))
    Fd = np.sum(2 * np.pi * wt * mu * (2.0 * (1 - t) - 2.0 * slope))
    assert abs(g["flux_up"][0] - Fu) < 1e-12 * Fu
    assert abs(g["flux_down"][1] - Fd) < 1e-12 * Fd


@pytest.mark.parametrize("nq", [8, 16, 32])
@pytest.mark.parametrize("delta_m,gg,mu0,tau", [(False, 0.7, 0.5, 0.3), (True, 0.85, 0.6, 0.5)])
def test_energy_conservation_on_gpu(nq, delta_m, gg, mu0, tau):  # test_ad_solver.cpp:1155-1203
    n = 5
    nm = 2 * nq + (1 if delta_m else 0)
    cfg = dict(num_layers=n, num_quadrature=nq, delta_tau=[tau] * n, single_scat_albedo=[1.0] * n,
               phase_function_moments=[list(gg ** np.arange(nm))] * n, use_delta_m=delta_m,
               solar_flux=1.0, solar_mu=mu0)
    g = _gpu_column(cfg)
    fnet = g["flux_up"] - g["flux_down"] - g["flux_direct"]
    assert np.all(np.abs(fnet - fnet[0]) < 1e-5)


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_delta_m_rayleigh_no_effect_on_gpu(nq):   # test_ad_solver.cpp:1210-1238
    ray = lambda nm: [1.0, 0.0, 0.1] + [0.0] * (nm - 3)
    base = dict(num_layers=1, num_quadrature=nq, delta_tau=[0.5], single_scat_albedo=[1.0],
                solar_flux=1.0, solar_mu=1.0)
    a = _gpu_column(dict(base, phase_function_moments=[ray(2 * nq)]))
    b = _gpu_column(dict(base, phase_function_moments=[ray(2 * nq + 1)], use_delta_m=True))
    for k in ("flux_up", "flux_down"):
        assert np.allclose(a[k], b[k], rtol=1e-12, atol=0)


def _thermal_column(nq, tau, ssa, B, surface_emission, top_emission, **bc):
    """One thermal column on the batched path, isotropic scattering."""
    L = len(tau)
    mom = np.zeros((L, 2 * nq))
    mom[:, 0] = 1.0
    cfg = BatchConfig(num_quadrature=nq, surface_emission=surface_emission,
                      top_emission=top_emission, **bc)
    out = solve_batch(_t(np.array(tau)[None]), _t(np.array(ssa)[None]), _t(mom), cfg,
                      planck_levels=_t(np.array(B)[None]))
    torch.cuda.synchronize()
    return {k: v[0].cpu().numpy() for k, v in out.items()}


_T4 = [200.0, 230.0, 260.0, 290.0]


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_top_emission_gives_pi_B_on_gpu(nq):   # test_ad_solver.cpp:1043-1084
    B = [oracle.planck(500.0, 600.0, T) for T in _T4]
    g = _thermal_column(nq, [0.5] * 3, [0.3] * 3, B, B[-1], B[0])
    assert abs(g["flux_down"][0] - np.pi * B[0]) <= 1e-12 * np.pi * B[0]
    g = _thermal_column(nq, [0.5] * 3, [0.3] * 3, B, B[-1], 0.0)   # cold space
    assert abs(g["flux_down"][0]) < 1e-12 * np.pi * B[0]


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_isothermal_divergence_vanishes_on_gpu(nq):   # test_ad_solver.cpp:1819-1840
    Bv = oracle.planck(300.0, 800.0, 255.0)
    g = _thermal_column(nq, [0.4] * 5, [0.5] * 5, [Bv] * 6, Bv, Bv)
    # the thin-layer start makes I = B a fixed point only to rounding amplified
    # by the doubling (oracle: |J / B - 1| <= 8.4e-12 at N = 32)
    assert np.all(np.abs(g["flux_divergence"]) <= 1e-10 * 4 * np.pi * Bv)
    assert np.allclose(g["mean_intensity"], Bv, rtol=1e-10, atol=0)


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_diffusion_bc_on_gpu(nq):   # test_ad_solver.cpp:1517-1558, 1589-1624
    B = [oracle.planck(500.0, 600.0, T) for T in _T4]
    g = _thermal_column(nq, [0.5] * 3, [0.0] * 3, B, 0.0, B[0], use_diffusion_lower_bc=True)
    # I_up(L) = B_L + mu dB/dtau: 2 pi sum w mu (B + mu dB) = pi B + (2 pi / 3) dB
    exp = np.pi * B[3] + (2 * np.pi / 3) * (B[3] - B[2]) / 0.5
    assert abs(g["flux_up"][3] - exp) <= 1e-12 * exp
    a = _thermal_column(nq, [0.5] * 3, [0.5] * 3, B, 0.0, B[0], use_diffusion_lower_bc=True)
    b = _thermal_column(nq, [0.5] * 3, [0.5] * 3, B, 0.0, B[0], use_diffusion_lower_bc=True,
                        surface_albedo=0.7)
    for k in OUTS:
        assert np.array_equal(a[k], b[k]), k


def _divergence_balance(g, dtau):
    """Trapezoid integral of the divergence over tau against the change of the
    net flux across the column (test_ad_solver.cpp:1842-1907)."""
    d = g["flux_divergence"]
    integral = np.sum(0.5 * (d[:-1] + d[1:]) * dtau)
    fnet = g["flux_up"] - g["flux_down"] - g["flux_direct"]
    return integral, fnet[-1] - fnet[0]


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_divergence_integral_thermal_on_gpu(nq):   # test_ad_solver.cpp:1842-1876
    n = 60
    B = [oracle.planck(300.0, 800.0, 200.0 + 1.5 * l) for l in range(n + 1)]
    g = _thermal_column(nq, [0.05] * n, [0.5] * n, B, B[-1], B[0])
    integral, dF = _divergence_balance(g, 0.05)
    assert abs(integral - dF) <= 0.01 * abs(dF)


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_divergence_integral_solar_on_gpu(nq):     # test_ad_solver.cpp:1878-1907
    n = 40
    g = _gpu_column(_iso_cfg(nq, [0.05] * n, [0.0] * n, solar_flux=200.0, solar_mu=0.5))
    integral, dF = _divergence_balance(g, 0.05)
    assert abs(integral - dF) <= 0.01 * abs(dF)


def _mixed_column_batch(nq, W=16, L=9, seed=3):
    rng = np.random.default_rng(seed)
    tau = 10 ** rng.uniform(-2, 0.7, (W, L))
    ssa = rng.uniform(0.05, 0.99, (W, L))
    mom = rng.uniform(0.3, 0.8, L)[:, None] ** np.arange(2 * nq)
    B = np.sort(rng.uniform(1.0, 3.0, (W, L + 1)), axis=1)
    return tau, ssa, mom, B


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_index_reversal_on_gpu(nq):   # test_ad_solver.cpp:1773-1812
    tau, ssa, mom, B = _mixed_column_batch(nq)
    cfg = dict(num_quadrature=nq, surface_albedo=0.2, solar_flux=1.0, solar_mu=0.6)
    a = solve_batch(_t(tau), _t(ssa), _t(mom), BatchConfig(**cfg), planck_levels=_t(B))
    b = solve_batch(_t(tau[:, ::-1]), _t(ssa[:, ::-1]), _t(mom[::-1]),
                    BatchConfig(index_from_bottom=True, **cfg), planck_levels=_t(B[:, ::-1]))
    torch.cuda.synchronize()
    for k in OUTS:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()[:, ::-1]
        assert np.allclose(x, y, rtol=1e-12, atol=1e-14 * np.abs(x).max()), k


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_workspace_reuse_on_gpu(nq):   # test_ad_solver.cpp:1733-1766
    tau, ssa, mom, B = _mixed_column_batch(nq)
    cfg = BatchConfig(num_quadrature=nq, surface_albedo=0.2, solar_flux=1.0, solar_mu=0.6)
    first = solve_batch(_t(tau), _t(ssa), _t(mom), cfg, planck_levels=_t(B))
    first = {k: v.cpu().numpy() for k, v in first.items()}
    other = solve_batch(_t(tau * 3.0), _t(ssa * 0.5), _t(mom), cfg, planck_levels=_t(B * 2.0))
    again = solve_batch(_t(tau), _t(ssa), _t(mom), cfg, planck_levels=_t(B))
    torch.cuda.synchronize()
    for k in OUTS:
        assert np.array_equal(first[k], again[k].cpu().numpy()), k
        assert not np.array_equal(first[k], other[k].cpu().numpy()) or k == "flux_direct", k


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_layer_split_invariance_on_gpu(nq):   # test_ad_solver.cpp:859-886
    kw = dict(solar_flux=1.0, solar_mu=0.5, top_emission=0.5 / np.pi, surface_albedo=0.5)
    a = _gpu_column(_iso_cfg(nq, [1.0], [0.9], **kw))
    b = _gpu_column(_iso_cfg(nq, [0.5, 0.5], [0.9, 0.9], **kw))
    for k in ("flux_up", "flux_down", "mean_intensity"):
        assert abs(a[k][0] - b[k][0]) < 1e-6 * abs(a[k][0]) + 1e-12, k
        assert abs(a[k][1] - b[k][2]) < 1e-6 * abs(a[k][1]) + 1e-12, k


def test_dynamic_quadrature_orders_converge_on_gpu():   # test_ad_solver.cpp:1959-1983, 2047-2075
    """Orders 6 and 12 run padded to 8 and 16 with zero-weight nodes."""
    vals = []
    for nq in (6, 8, 12, 16):
        g = _gpu_column(dict(num_layers=1, num_quadrature=nq, delta_tau=[0.5],
                             single_scat_albedo=[0.9],
                             phase_function_moments=[list(0.5 ** np.arange(2 * nq))],
                             solar_flux=1.0, solar_mu=0.5))
        ref = oracle.solve(num_layers=1, num_quadrature=nq, delta_tau=[0.5],
                           single_scat_albedo=[0.9],
                           phase_function_moments=[list(0.5 ** np.arange(2 * nq))],
                           solar_flux=1.0, solar_mu=0.5)
        assert abs(g["flux_up"][0] - ref["flux_up"][0]) <= 1e-10 * abs(ref["flux_up"][0]), nq
        vals.append(g["flux_up"][0])
    assert abs(vals[0] - vals[-1]) < 0.01 * abs(vals[-1])
    assert abs(vals[2] - vals[3]) < abs(vals[0] - vals[3])
