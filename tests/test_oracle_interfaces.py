"""The oracle's adding sweeps and interface solves against a direct solve.

The adding method with interface solves (src/adding.hpp:31-258,
src/solver.cpp:24-73,560-700 of the reference; oracle/adrt_oracle.c
add_general, add_nonscat_top, add_nonscat_bot, interface_up, interface_down)
is a direct solver of the block two-stream system of the interface
intensities.  For layers l = 0..L-1 with records (R_l, T_l, s_up_l, s_dn_l)
(symmetric layers: R_ab = R_ba, T_ab = T_ba):

    I_up(l)   = R_l I_dn(l) + T_l I_up(l+1) + s_up_l
    I_dn(l+1) = T_l I_dn(l) + R_l I_up(l+1) + s_dn_l

with I_dn(0) the top emission and, at the bottom, I_up(L) = R_s I_dn(L) + s_s
for the Lambertian surface layer or the given I_up(L) (diffusion boundary or
no surface).  Here that system is assembled from the oracle's per-layer
doubling records and solved as one dense linear system in numpy; the fluxes
and mean intensity it gives must equal the oracle column solve to 1e-12 of
each column's largest value.

The columns mix scattering layers with non-scattering (omega = 0) and
transparent (tau = 0) layers above and below them, so every dispatch branch
of the adding runs: a non-scattering layer on top of a scattering composite
and below one, non-scattering composites, a non-scattering surface (albedo 0,
emitting), and bottom boundaries with a non-zero transmission (diffusion lower
boundary).  Composites of two or more different layers are not symmetric, so
an exchange of T_ab and T_ba (or R_ab and R_ba) in any variant changes the
result; the last test shows that the check resolves such an exchange.
"""
import math

import numpy as np
import pytest

import oracle

TOL = 1e-12


def moments(g, M):
    return g ** np.arange(M)


def layer_records(n, tau, omega, g, B, F0, mu0, M):
    """Per-layer R, T and combined sources from the oracle's doubling, with the
    phase matrices and solar vectors of the oracle (pinned elsewhere)."""
    mu, wt = oracle.gauss_legendre(n)
    recs = []
    tau_cum = 0.0
    for l in range(len(tau)):
        chi = moments(g[l], M)
        if omega[l] > 0 and tau[l] > 0:
            Ppp, Ppm = oracle.phase_matrices(chi, mu, wt)
            pp, pm = oracle.solar_phase_vectors(chi, mu, wt, mu0) if F0 > 0 else (None, None)
        else:
            Ppp = Ppm = np.zeros((n, n))
            pp = pm = None
        d = oracle.doubling(n, tau[l], omega[l], B[l], B[l + 1], Ppp, Ppm, mu, wt,
                            solar_flux=F0, solar_mu=mu0, tau_cumulative=tau_cum,
                            p_plus=pp if F0 > 0 else None, p_minus=pm if F0 > 0 else None)
        su = d["s_up"] + (d["s_up_solar"] if F0 > 0 else 0.0)
        sd = d["s_down"] + (d["s_down_solar"] if F0 > 0 else 0.0)
        recs.append((d["R"], d["T"], su, sd))
        tau_cum += tau[l]
    return recs, mu, wt, tau_cum


def direct_solve(n, recs, mu, wt, Itd, bottom, transpose_T_of=None):
    """Dense solve of the interface system.  bottom = ("surface", R_s, s_s) or
    ("given", I_up(L)).  transpose_T_of: a layer whose T enters transposed (a
    planted orientation mistake, for the sensitivity check)."""
    L = len(recs)
    nu = 2 * n * (L + 1)
    up = lambda l: slice(2 * n * l, 2 * n * l + n)
    dn = lambda l: slice(2 * n * l + n, 2 * n * (l + 1))
    A = np.zeros((nu, nu))
    b = np.zeros(nu)
    I = np.eye(n)
    row = 0

    def eq():
        nonlocal row
        r = slice(row, row + n)
        row += n
        return r

    r = eq()
    A[r, dn(0)] = I
    b[r] = Itd
    for l, (R, T, su, sd) in enumerate(recs):
        if transpose_T_of == l:
            T = T.T
        r = eq()  # I_up(l) - R I_dn(l) - T I_up(l+1) = s_up
        A[r, up(l)] = I
        A[r, dn(l)] -= R
        A[r, up(l + 1)] -= T
        b[r] = su
        r = eq()  # I_dn(l+1) - T I_dn(l) - R I_up(l+1) = s_dn
        A[r, dn(l + 1)] = I
        A[r, dn(l)] -= T
        A[r, up(l + 1)] -= R
        b[r] = sd
    r = eq()
    A[r, up(L)] = I
    if bottom[0] == "surface":
        A[r, dn(L)] -= bottom[1]
        b[r] = bottom[2]
    else:
        b[r] = bottom[1]
    x = np.linalg.solve(A, b)
    Iup = np.array([x[up(l)] for l in range(L + 1)])
    Idn = np.array([x[dn(l)] for l in range(L + 1)])
    c = 2.0 * math.pi * wt * mu
    return Iup @ c, Idn @ c, 0.5 * (Iup + Idn) @ wt


CASES = {
    # name: (tau, omega, surface albedo, surface emission, diffusion, solar)
    "nonscat_between_scat_surface": ([0.4, 0.7, 1.3, 0.2], [0.8, 0.0, 0.95, 0.5], 0.3, 1.1,
                                     False, True),
    "nonscat_top_and_bottom_emitting_black_surface": ([0.6, 0.9, 0.3, 1.7, 0.5],
                                                      [0.0, 0.7, 0.0, 0.9, 0.0], 0.0, 1.4,
                                                      False, True),
    "nonscat_below_scat_diffusion": ([0.5, 1.1, 0.8, 0.6], [0.9, 0.6, 0.0, 0.0], 0.2, 0.0,
                                     True, False),
    "nonscat_above_scat_diffusion": ([0.9, 0.4, 2.0, 0.7], [0.0, 0.0, 0.99, 0.3], 0.0, 0.0,
                                     True, True),
    "transparent_and_nonscat_mixed": ([0.5, 0.0, 0.8, 0.0, 0.3, 1.2], [0.7, 0.4, 0.0, 0.9, 0.6, 0.0],
                                      0.45, 0.9, False, True),
    "alternating_diffusion": ([0.3, 0.8, 0.5, 1.5, 0.4], [0.95, 0.0, 0.85, 0.0, 0.75], 0.0, 0.0,
                              True, True),
}


def _column(n, case, seed=0):
    tau, omega, A, Bs, diffusion, solar = CASES[case]
    tau, omega = np.array(tau, float), np.array(omega, float)
    L = len(tau)
    rng = np.random.default_rng(seed + n)
    g = rng.uniform(0.2, 0.85, L)
    B = np.linspace(1.0, 2.2, L + 1) + rng.uniform(0, 0.2, L + 1)
    Bt = 0.35
    F0, mu0 = (1.7, 0.55) if solar else (0.0, 1.0)
    return tau, omega, g, B, A, Bs, Bt, diffusion, F0, mu0


def _oracle_column(n, tau, omega, g, B, A, Bs, Bt, diffusion, F0, mu0, M):
    mom = np.array([moments(gl, M) for gl in g])
    return oracle.solve_batch(tau[None], omega[None], mom[None], n, planck_levels=B[None],
                              surface_emission=Bs, top_emission=Bt, surface_albedo=A,
                              solar_flux=F0, solar_mu=mu0, use_diffusion_lower_bc=diffusion)


def _bottom(n, recs, mu, wt, tau, B, A, Bs, diffusion, F0, mu0, tau_tot):
    if diffusion:
        dB = (B[-1] - B[-2]) / tau[-1] if tau[-1] > 0 else 0.0
        return ("given", B[-1] + mu * dB)
    if A > 0 or Bs > 0:
        xfac = 0.5 / np.sum(mu * wt)
        Rs = np.tile(2.0 * A * mu * wt * xfac, (n, 1))
        ss = (1.0 - A) * Bs * np.ones(n)
        if F0 > 0 and A > 0:
            ss = ss + (A / math.pi) * F0 * mu0 * math.exp(-tau_tot / mu0)
        return ("surface", Rs, ss)
    return ("given", Bs * np.ones(n))


@pytest.mark.parametrize("n", [4, 8])
@pytest.mark.parametrize("case", sorted(CASES))
def test_adding_and_interfaces_match_direct_solve(n, case):
    tau, omega, g, B, A, Bs, Bt, diffusion, F0, mu0 = _column(n, case)
    M = 2 * n
    recs, mu, wt, tau_tot = layer_records(n, tau, omega, g, B, F0, mu0, M)
    fu, fd, J = direct_solve(n, recs, mu, wt, Bt * np.ones(n),
                             _bottom(n, recs, mu, wt, tau, B, A, Bs, diffusion, F0, mu0, tau_tot))
    r = _oracle_column(n, tau, omega, g, B, A, Bs, Bt, diffusion, F0, mu0, M)
    fdir = np.zeros(len(tau) + 1)
    if F0 > 0:
        fdir = F0 * mu0 * np.exp(-np.concatenate([[0.0], np.cumsum(tau)]) / mu0)
        J = J + fdir / (4.0 * math.pi * mu0)
    for name, ref in (("flux_up", fu), ("flux_down", fd), ("mean_intensity", J),
                      ("flux_direct", fdir)):
        got = r[name][0]
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(got - ref).max() / scale
        assert err <= TOL, (case, n, name, err)


@pytest.mark.parametrize("case", ["nonscat_between_scat_surface", "nonscat_below_scat_diffusion",
                                  "alternating_diffusion"])
def test_direct_solve_resolves_a_transposed_transmission(case):
    """A check with teeth: entering one scattering layer's T transposed (the
    mistake an exchange of T_ab and T_ba would make) moves the direct solution
    by far more than the tolerance of the comparison above."""
    n = 8
    tau, omega, g, B, A, Bs, Bt, diffusion, F0, mu0 = _column(n, case)
    recs, mu, wt, tau_tot = layer_records(n, tau, omega, g, B, F0, mu0, 2 * n)
    bottom = _bottom(n, recs, mu, wt, tau, B, A, Bs, diffusion, F0, mu0, tau_tot)
    good = direct_solve(n, recs, mu, wt, Bt * np.ones(n), bottom)
    lyr = int(np.flatnonzero(omega > 0)[0])
    bad = direct_solve(n, recs, mu, wt, Bt * np.ones(n), bottom, transpose_T_of=lyr)
    diff = max(np.abs(a - b).max() / np.abs(a).max() for a, b in zip(good, bad))
    assert diff > 1e4 * TOL, diff
