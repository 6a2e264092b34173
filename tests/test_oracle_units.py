"""Pin the oracle's building blocks against library routines and exact math.

* quadrature  -> numpy.polynomial.legendre.leggauss mapped to [0, 1]
* Legendre    -> scipy.special.eval_legendre
* phase       -> brute-force numpy of the addition-theorem sum + Hansen invariant
* Planck      -> scipy.integrate.quad of the Planck integrand, Stefan-Boltzmann,
                 central differences for dB/dT
* doubling    -> the exact layer operator from the matrix exponential of the
                 discrete-ordinate generator (R, T, thermal and solar sources),
                 to the first-order-init discretisation error (~1e-5)
"""
import math

import numpy as np
import pytest
import scipy.integrate
import scipy.linalg
import scipy.special

import oracle
from golden.reference_tables import moments

PI = math.pi


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16, 32])
def test_quadrature_matches_leggauss(n):
    mu, wt = oracle.gauss_legendre(n)
    x, w = np.polynomial.legendre.leggauss(n)
    assert np.allclose(mu, np.sort((1 + x) / 2), rtol=0, atol=1e-14)
    assert np.allclose(wt, (w / 2)[np.argsort((1 + x) / 2)], rtol=0, atol=1e-14)
    assert abs(wt.sum() - 1.0) < 1e-14
    assert abs((mu * wt).sum() - 0.5) < 1e-14
    assert np.all((mu > 0) & (mu < 1)) and np.all(np.diff(mu) > 0)


def test_legendre_table_matches_scipy():
    x = np.linspace(-1, 1, 9)
    P = oracle.legendre_table(40, x)
    for l in range(40):
        assert np.allclose(P[l], scipy.special.eval_legendre(l, x), atol=1e-13)


def _phase_bruteforce(chi, mu, wt):
    L = len(chi)
    Pl = np.array([scipy.special.eval_legendre(l, mu) for l in range(L)])
    c = (2 * np.arange(L) + 1) * np.asarray(chi)
    s = (-1.0) ** np.arange(L)
    Ppp = np.einsum("l,li,lj->ij", c, Pl, Pl) / (2 * PI)
    Ppm = np.einsum("l,li,lj->ij", c * s, Pl, Pl) / (2 * PI)
    col = ((Ppp + Ppm) * wt[:, None]).sum(0)
    corr = np.where(col > 0, 1.0 / (2 * PI * np.where(col > 0, col, 1.0)), 1.0)
    return Ppp * corr[None, :], Ppm * corr[None, :]


@pytest.mark.parametrize("kind,kw", [("isotropic", {}), ("rayleigh", {}), ("hg", {"g": 0.8}),
                                     ("haze", {}), ("dhg", {"g": 0.7, "f": 0.8, "g2": -0.3})])
@pytest.mark.parametrize("n", [4, 8, 16])
def test_phase_matrices(kind, kw, n):
    mu, wt = oracle.gauss_legendre(n)
    chi = moments(kind, 2 * n, **kw)
    Ppp, Ppm = oracle.phase_matrices(chi, mu, wt)
    Bpp, Bpm = _phase_bruteforce(chi, mu, wt)
    assert np.allclose(Ppp, Bpp, rtol=1e-12, atol=1e-14)
    assert np.allclose(Ppm, Bpm, rtol=1e-12, atol=1e-14)
    # Hansen normalisation: every column integrates to 1/(2 pi)
    col = ((Ppp + Ppm) * wt[:, None]).sum(0)
    assert np.allclose(col, 1.0 / (2 * PI), rtol=1e-13)
    if kind == "isotropic":
        assert np.allclose(Ppp, 1 / (4 * PI)) and np.allclose(Ppm, 1 / (4 * PI))


def test_solar_phase_vectors():
    n = 8
    mu, wt = oracle.gauss_legendre(n)
    chi = moments("hg", 16, g=0.6)
    mu0 = 0.37
    pp, pm = oracle.solar_phase_vectors(chi, mu, wt, mu0)
    L = len(chi)
    c = (2 * np.arange(L) + 1) * np.asarray(chi)
    Pl = np.array([scipy.special.eval_legendre(l, mu) for l in range(L)])
    P0 = scipy.special.eval_legendre(np.arange(L), mu0)
    rp = (c * P0) @ Pl / (2 * PI)
    rm = (c * P0 * (-1.0) ** np.arange(L)) @ Pl / (2 * PI)
    k = 1.0 / (2 * PI * ((rp + rm) * wt).sum())   # single scalar Hansen factor
    assert np.allclose(pp, rp * k, rtol=1e-12)
    assert np.allclose(pm, rm * k, rtol=1e-12)


def _planck_quad(lo, hi, T):
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    # the reference constants: C2 = hc/k in cm K, sigma as in planck.cpp
    C2 = 1.438786
    sigma = 5.67032e-8
    conc = 15.0 / PI ** 4
    f = lambda v: v ** 3 / np.expm1(v)
    val, _ = scipy.integrate.quad(f, C2 * lo / T, C2 * hi / T, epsabs=0, epsrel=1e-12,
                                  limit=200)
    return sigma / PI * T ** 4 * conc * val


@pytest.mark.parametrize("lo,hi,T", [(500, 600, 300), (300, 800, 200), (0.01, 50000, 300),
                                     (1000, 1005, 250), (2000, 2500, 180), (10, 30, 310)])
def test_planck_matches_quadrature(lo, hi, T):
    # The DisORT series (src/planck.cpp:83-130) truncates at ~1e-5 relative for
    # bands where d(v0) - d(v1) cancels; the narrow-band Simpson branch is ~1e-6.
    B = oracle.planck(lo, hi, T)
    ref = _planck_quad(lo, hi, T)
    assert abs(B - ref) <= 1e-4 * ref


def test_planck_stefan_boltzmann_and_edges():   # test_ad_solver.cpp:116-136
    B = oracle.planck(0.01, 100000.0, 300.0)
    exp = 5.670374419e-8 * 300.0 ** 4 / PI
    assert abs(B - exp) < 0.01 * exp
    assert oracle.planck(500, 600, 0.0) == 0.0
    assert oracle.planck(500, 600, 400) > oracle.planck(500, 600, 300) > 0
    with pytest.raises(oracle.OracleError):
        oracle.planck(600, 500, 300)


@pytest.mark.parametrize("lo,hi,T", [(500, 600, 300), (300, 800, 200), (0.01, 50000, 250),
                                     (1000, 1005, 250), (700, 700, 260)])
def test_planck_derivative_fd(lo, hi, T):
    d = oracle.planck_deriv(lo, hi, T)
    h = 1e-3
    fd = (oracle.planck(lo, hi, T + h) - oracle.planck(lo, hi, T - h)) / (2 * h)
    assert abs(d - fd) <= 1e-5 * abs(fd)


def test_doubling_count():               # src/doubling.hpp:20-26,107-109
    assert oracle.doubling_count(1.0, 0.5) == 16
    assert oracle.doubling_count(1.0, 0.05) == 10
    assert oracle.doubling_count(1.0, 0.005) == 4
    assert oracle.doubling_count(0.5, 0.5) == 15
    assert oracle.doubling_count(0.4, 0.5) == 15          # trunc(-1.32) = -1
    assert oracle.doubling_count(8.0, 0.5) == 19
    assert oracle.doubling_count(1e-3, 0.005) == 1        # trunc(-9.97)+4 < 1 -> 1
    assert oracle.doubling_count(1.0, 1.5) == 16          # omega clamped


def _exact_layer(n, tau, omega, chi, B_top, B_bot, F_top=0.0, mu0=0.5):
    """Exact discrete-ordinate layer operator via the matrix exponential."""
    mu, wt = oracle.gauss_legendre(n)
    Ppp, Ppm = oracle.phase_matrices(chi, mu, wt)
    c = 2 * PI * omega
    Gpp = (np.eye(n) - c * Ppp * wt[None, :]) / mu[:, None]
    Gpm = c * Ppm * wt[None, :] / mu[:, None]
    pp, pm = oracle.solar_phase_vectors(chi, mu, wt, mu0)
    Bd = (B_bot - B_top) / tau
    # state [I_dn, I_up, c0 (=1), c1 (=tau'), e (=exp(-tau'/mu0))]
    m = 2 * n + 3
    H = np.zeros((m, m))
    H[:n, :n] = -Gpp
    H[:n, n:2 * n] = Gpm
    H[n:2 * n, :n] = -Gpm
    H[n:2 * n, n:2 * n] = Gpp
    src = (1 - omega) / mu
    H[:n, 2 * n] = src * B_top
    H[:n, 2 * n + 1] = src * Bd
    H[n:2 * n, 2 * n] = -src * B_top
    H[n:2 * n, 2 * n + 1] = -src * Bd
    H[:n, 2 * n + 2] = omega * F_top * pp / mu
    H[n:2 * n, 2 * n + 2] = -omega * F_top * pm / mu
    H[2 * n + 1, 2 * n] = 1.0
    H[2 * n + 2, 2 * n + 2] = -1.0 / mu0
    Mx = scipy.linalg.expm(tau * H)
    M11, M12 = Mx[:n, :n], Mx[:n, n:2 * n]
    M21, M22 = Mx[n:2 * n, :n], Mx[n:2 * n, n:2 * n]
    Ms1 = Mx[:n, 2 * n:]
    Ms2 = Mx[n:2 * n, 2 * n:]
    z0 = np.array([1.0, 0.0, 1.0])
    M22i = np.linalg.inv(M22)
    R = -M22i @ M21
    T = M22i
    # zero incident radiation: I_up(0) solves I_up(tau)=0
    Iup0 = -M22i @ (Ms2 @ z0)
    Idn_tau = M12 @ Iup0 + Ms1 @ z0
    return R, T, Iup0, Idn_tau, (mu, wt, Ppp, Ppm, pp, pm)


def _star(top, bot):
    """Independent numpy star product of two layers (interaction principle)."""
    Ra, Rb, Ta, Tb, ua, da = top
    Rc, Rd, Tc, Td, uc, dc = bot   # R_ab, R_ba, T_ab, T_ba, s_up, s_down
    n = Ra.shape[0]
    I = np.eye(n)
    D1 = np.linalg.inv(I - Rc @ Rb)
    D2 = np.linalg.inv(I - Rb @ Rc)
    R_ab = Ra + Tb @ D1 @ Rc @ Ta
    R_ba = Rd + Tc @ D2 @ Rb @ Td
    T_ab = Tc @ D2 @ Ta
    T_ba = Tb @ D1 @ Td
    s_up = ua + Tb @ D1 @ (uc + Rc @ da)
    s_dn = dc + Tc @ D2 @ (da + Rb @ uc)
    return R_ab, R_ba, T_ab, T_ba, s_up, s_dn


def _exact_thick(n, tau, omega, chi, B_top, B_bot, F_top, mu0, k=6):
    """Exact operator of a thick layer: 2^k exact (expm) sub-layers, star-added."""
    sub = tau / 2 ** k
    layers = []
    for j in range(2 ** k):
        bt = B_top + (B_bot - B_top) * j / 2 ** k
        bb = B_top + (B_bot - B_top) * (j + 1) / 2 ** k
        ft = F_top * math.exp(-j * sub / mu0)
        R, T, u, d, aux = _exact_layer(n, sub, omega, chi, bt, bb, ft, mu0)
        layers.append((R, R, T, T, u, d))
    while len(layers) > 1:
        layers = [_star(layers[i], layers[i + 1]) for i in range(0, len(layers), 2)]
    R_ab, R_ba, T_ab, T_ba, u, d = layers[0]
    assert np.allclose(R_ab, R_ba, atol=1e-12) and np.allclose(T_ab, T_ba, atol=1e-12)
    return R_ab, T_ab, u, d, aux


@pytest.mark.parametrize("n,tau,omega,kind,g", [
    (4, 0.3, 0.9, "hg", 0.6), (8, 0.05, 0.5, "rayleigh", 0), (2, 1.0, 0.99, "isotropic", 0),
    (8, 0.08, 0.2, "hg", 0.3), (8, 2.5, 0.95, "hg", 0.7), (4, 6.0, 0.6, "rayleigh", 0)])
def test_doubling_matches_exact_layer(n, tau, omega, kind, g):
    chi = moments(kind, 2 * n, g=g)
    B_top, B_bot, F_top, mu0 = 1.3, 2.1, 0.7, 0.45
    if tau / oracle.gauss_legendre(n)[0][0] < 6.0:
        R, T, Iup0, Idn, aux = _exact_layer(n, tau, omega, chi, B_top, B_bot, F_top, mu0)
    else:
        R, T, Iup0, Idn, aux = _exact_thick(n, tau, omega, chi, B_top, B_bot, F_top, mu0)
    mu, wt, Ppp, Ppm, pp, pm = aux
    # tau_cumulative chosen so that F_top = solar_flux * exp(-tau_cum/mu0)
    tau_cum = 0.2
    flux = F_top * math.exp(tau_cum / mu0)
    d = oracle.doubling(n, tau, omega, B_top, B_bot, Ppp, Ppm, mu, wt, solar_flux=flux,
                        solar_mu=mu0, tau_cumulative=tau_cum, p_plus=pp, p_minus=pm)
    assert d["is_scattering"]
    tol = 2e-4
    assert np.max(np.abs(d["R"] - R)) < tol * np.max(np.abs(R))
    assert np.max(np.abs(d["T"] - T)) < tol * np.max(np.abs(T))
    up = d["s_up"] + d["s_up_solar"]
    dn = d["s_down"] + d["s_down_solar"]
    assert np.max(np.abs(up - Iup0)) < tol * np.abs(Iup0).max()
    assert np.max(np.abs(dn - Idn)) < tol * np.abs(Idn).max()


def test_pure_absorption_layer_closed_form():   # src/doubling.hpp:56-77
    n, tau = 8, 0.8
    mu, wt = oracle.gauss_legendre(n)
    Z = np.zeros((n, n))
    d = oracle.doubling(n, tau, 0.0, 1.0, 3.0, Z, Z, mu, wt)
    t = np.exp(-tau / mu)
    assert not d["is_scattering"]
    assert np.allclose(np.diag(d["T"]), t, rtol=1e-15) and np.all(d["R"] == 0)
    Bb, Bd = 2.0, 2.0 / tau
    slope = mu * (1 - t) - 0.5 * tau * (1 + t)
    assert np.allclose(d["s_up"], Bb * (1 - t) + Bd * slope, rtol=1e-14)
    assert np.allclose(d["s_down"], Bb * (1 - t) - Bd * slope, rtol=1e-14)


def test_transparent_layer():
    n = 4
    mu, wt = oracle.gauss_legendre(n)
    Z = np.zeros((n, n))
    d = oracle.doubling(n, 0.0, 0.5, 1.0, 3.0, Z, Z, mu, wt)
    assert not d["is_scattering"]
    assert np.all(d["T"] == np.eye(n)) and np.all(d["R"] == 0) and np.all(d["s_up"] == 0)


def test_conservative_layer_flux_conservation():
    """omega = 1: reflected + transmitted flux equals incident (sum_i mu_i w_i)."""
    n = 8
    mu, wt = oracle.gauss_legendre(n)
    chi = moments("hg", 16, g=0.7)
    Ppp, Ppm = oracle.phase_matrices(chi, mu, wt)
    d = oracle.doubling(n, 1.3, 1.0, 0, 0, Ppp, Ppm, mu, wt)
    mw = mu * wt
    out = mw @ (d["R"] + d["T"])
    assert np.allclose(out, mw, rtol=2e-6)
