"""Pin the oracle's raw-coefficient restatement (oracle.raw_optics,
oracle.planck_single) against constants, library routines and exact math.

* planck_single -> CODATA constants (scipy.constants) for c1 = 2 h c^2 and
  c2 = h c / k in the units of src_cuda/cuda_planck_single.cuh; its integral
  over a band (scipy.integrate.quad) against the oracle's band Planck
  function, itself pinned by test_oracle_units; the T < 1e-5 K cut-off
* raw_optics    -> column optical depth and scattering depth against
  numpy.trapezoid over altitude (a wrong level pairing, a dropped half or a
  missing layer reversal changes the per-layer values these pin); a linear
  profile whose layer integral is exact; a single absorbing level that must
  land in the two layers next to it, TOA first; cloud addition and the clamp
  at 0 (src_cuda/cuda_solver.cu:1131-1182); the surface / top temperature
  rules (src_cuda/cuda_solver.cu:220-241,428-441)
"""
import numpy as np
import pytest
import scipy.constants as sc
import scipy.integrate

import oracle


def test_planck_single_constants_match_codata():
    # B_nu per cm^-1 in W m^-2 cm: 2 h c^2 nu^3 with nu in m^-1 = 100 nu_cm and
    # a factor 100 from per-m^-1 to per-cm^-1
    c1 = 2 * sc.h * sc.c ** 2 * 1e8
    c2 = sc.h * sc.c / sc.k * 100.0
    assert abs(oracle.PLANCK_SINGLE_C1 - c1) <= 1e-5 * c1
    assert abs(oracle.PLANCK_SINGLE_C2 - c2) <= 2e-5 * c2


@pytest.mark.parametrize("lo,hi,T", [(500, 600, 300), (1000, 1100, 250), (2000, 2200, 320),
                                     (20, 40, 180)])
def test_planck_single_integrates_to_band_planck(lo, hi, T):
    val, _ = scipy.integrate.quad(lambda v: float(oracle.planck_single(T, v)), lo, hi,
                                  epsabs=0, epsrel=1e-12)
    ref = oracle.planck(lo, hi, T)
    assert abs(val - ref) <= 1e-4 * ref


def test_planck_single_cutoff_and_monotone():
    nu = np.array([500.0, 1500.0])
    assert np.all(oracle.planck_single(0.0, nu) == 0.0)
    assert np.all(oracle.planck_single(9e-6, nu) == 0.0)
    assert np.all(oracle.planck_single(1e-5, nu) >= 0.0)
    b1, b2 = oracle.planck_single(250.0, nu), oracle.planck_single(300.0, nu)
    assert np.all(b2 > b1) and np.all(b1 > 0)


def _profile(L=9, W=5, seed=1):
    rng = np.random.default_rng(seed)
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(1e4, 5e5, L))])   # cm, BOA first
    a = 10 ** rng.uniform(-9, -5, (L + 1, W))
    s = 10 ** rng.uniform(-9, -6, (L + 1, W))
    T = np.linspace(310.0, 190.0, L + 1)
    nu = np.linspace(400.0, 2500.0, W)
    return z, a, s, T, nu


def test_column_depths_match_trapezoid():
    z, a, s, T, nu = _profile()
    r = oracle.raw_optics(a, s, z, T, nu)
    tau, ssa = r["delta_tau"], r["single_scat_albedo"]
    W, L = tau.shape
    for w in range(W):
        col = np.trapezoid(a[:, w] + s[:, w], z)
        assert abs(tau[w].sum() - col) <= 1e-13 * col
        sd = np.trapezoid(s[:, w], z)
        assert abs((ssa[w] * tau[w]).sum() - sd) <= 1e-13 * sd
        # layer by layer, TOA first: layer l spans BOA levels L-1-l .. L-l
        for l in range(L):
            b = L - 1 - l
            lay = np.trapezoid(a[b:b + 2, w] + s[b:b + 2, w], z[b:b + 2])
            assert abs(tau[w, l] - lay) <= 1e-15 * lay + 1e-300


def test_linear_profile_layer_integral_is_exact():
    L, W = 6, 3
    z = np.array([0.0, 1.0, 3.0, 4.0, 7.0, 8.0, 10.0]) * 1e5
    k0, k1 = np.array([1e-6, 2e-6, 5e-7]), np.array([-5e-13, -1e-12, -2e-13])
    a = k0[None, :] + k1[None, :] * z[:, None]
    s = 0.25 * a
    r = oracle.raw_optics(a, s, z, np.full(L + 1, 250.0), np.array([500.0, 600.0, 700.0]))
    for l in range(L):
        b = L - 1 - l
        exact = 1.25 * (k0 * (z[b + 1] - z[b]) + 0.5 * k1 * (z[b + 1] ** 2 - z[b] ** 2))
        assert np.allclose(r["delta_tau"][:, l], exact, rtol=1e-13, atol=0)
    assert np.allclose(r["single_scat_albedo"], 0.2, rtol=1e-14, atol=0)


def test_single_level_lands_in_adjacent_layers_toa_first():
    L, W = 5, 2
    z = np.arange(L + 1) * 1e5
    a = np.zeros((L + 1, W))
    a[4] = [1e-5, 2e-5]            # BOA level 4 = TOA level 1
    r = oracle.raw_optics(a, np.zeros_like(a), z, np.full(L + 1, 200.0), np.array([1.0, 2.0]))
    tau = r["delta_tau"]
    assert np.all(tau[:, [2, 3, 4]] == 0.0)
    assert np.allclose(tau[:, 0], 0.5 * a[4] * 1e5) and np.allclose(tau[:, 1], 0.5 * a[4] * 1e5)
    assert np.all(r["single_scat_albedo"] == 0.0)


def test_cloud_depth_and_clamp():
    z, a, s, T, nu = _profile(L=4, W=3)
    L, W = 4, 3
    base = oracle.raw_optics(a, s, z, T, nu)
    cloud = np.zeros((L, W))
    cloud[0, 0] = 0.7              # bottom layer = TOA-first index L-1
    cloud[2, 1] = -1e9             # drives the layer negative: clamped, omega 0
    r = oracle.raw_optics(a, s, z, T, nu, cloud_optical_depth=cloud)
    assert r["delta_tau"][0, L - 1] == base["delta_tau"][0, L - 1] + 0.7
    sd = base["single_scat_albedo"][0, L - 1] * base["delta_tau"][0, L - 1]
    assert abs(r["single_scat_albedo"][0, L - 1] * r["delta_tau"][0, L - 1] - sd) <= 1e-15
    assert r["delta_tau"][1, L - 3] == 0.0 and r["single_scat_albedo"][1, L - 3] == 0.0
    keep = np.ones((W, L), bool)
    keep[0, L - 1] = keep[1, L - 3] = False
    assert np.array_equal(r["delta_tau"][keep], base["delta_tau"][keep])


def test_boundary_temperature_rules():
    z, a, s, T, nu = _profile(L=3, W=4)
    r = oracle.raw_optics(a, s, z, T, nu)
    assert np.array_equal(r["planck_levels"][:, 0], oracle.planck_single(T[-1], nu))
    assert np.array_equal(r["planck_levels"][:, -1], oracle.planck_single(T[0], nu))
    assert np.array_equal(r["surface_emission"], r["planck_levels"][:, -1])
    assert np.array_equal(r["top_emission"], r["planck_levels"][:, 0])
    r = oracle.raw_optics(a, s, z, T, nu, surface_temperature=330.0, top_temperature=0.0)
    assert np.array_equal(r["surface_emission"], oracle.planck_single(330.0, nu))
    assert np.all(r["top_emission"] == 0.0)
