"""The adding algorithm of the HIP kernels, restated in numpy, against the
direct block solve and the oracle's adding.

The adding kernels (adding16_kernel, adding32_kernel, adding_kernel<N>) do not
build the reference's composites of the layers above and below every
interface (src/adding.hpp:69-258, src/solver.cpp:24-73,560-700).  They solve
the same two-stream system of interface intensities (see
test_oracle_interfaces.py) by one sweep of base composites from the bottom,

    X_l = (I - R_l Rb(l+1))^-1,   G_l = X_l T_l,   d_l = X_l (R_l sb(l+1) + s_dn(l)),
    Rb(l) = R_l + T_l Rb(l+1) G_l,   sb(l) = s_up(l) + T_l (Rb(l+1) d_l + sb(l+1)),

from Rb(L) = R_s, sb(L) = s_s (surface) or Rb(L) = 0, sb(L) = the given
I_up(L), and a downward sweep I_dn(l+1) = G_l I_dn(l) + d_l,
I_up(l) = Rb(l) I_dn(l) + sb(l).  Checked here:

* on the mixed scattering / non-scattering / transparent columns of
  test_oracle_interfaces.py it equals the direct solve to 1e-12;
* where two consecutive layers have an unstable doubling start (omega < 0.01,
  tau0 above the smallest quadrature cosine: R and T come out with entries far
  above one), it lands much further from the direct solve than the
  reference's adding does -- the reason the kernels route such columns
  (the prep kernel's per-column flag) to a kernel of the reference's
  composite form;
* with one ordinary scattering layer between the two unstable layers it is
  as close to the direct solve as the reference's adding, so the flag asks
  for consecutive unstable layers only.
"""
import numpy as np
import pytest

import oracle  # noqa: F401  (the records come from the oracle's doubling)
import test_oracle_interfaces as TI

TOL = 1e-12


def base_sweep(n, recs, Itd, bottom):
    """Interface intensities by the base sweep and the downward sweep."""
    L = len(recs)
    if bottom[0] == "surface":
        Rb, sb = bottom[1].copy(), bottom[2].copy()
    else:
        Rb, sb = np.zeros((n, n)), np.asarray(bottom[1], float).copy()
    G, d, Rbs, sbs = [None] * L, [None] * L, [None] * (L + 1), [None] * (L + 1)
    Rbs[L], sbs[L] = Rb, sb
    for l in range(L - 1, -1, -1):
        R, T, su, sd = recs[l]
        M = np.eye(n) - R @ Rb
        G[l] = np.linalg.solve(M, T)
        d[l] = np.linalg.solve(M, R @ sb + sd)
        Rb, sb = R + T @ Rb @ G[l], su + T @ (Rb @ d[l] + sb)
        Rbs[l], sbs[l] = Rb, sb
    Idn, Iup = [Itd], [Rbs[0] @ Itd + sbs[0]]
    for l in range(L):
        Idn.append(G[l] @ Idn[-1] + d[l])
        Iup.append(Rbs[l + 1] @ Idn[-1] + sbs[l + 1])
    return np.array(Iup), np.array(Idn)


def _fluxes(n, wt, mu, Iup, Idn):
    c = 2.0 * np.pi * wt * mu
    return Iup @ c, Idn @ c, 0.5 * (Iup + Idn) @ wt


@pytest.mark.parametrize("n", [4, 8])
@pytest.mark.parametrize("case", sorted(TI.CASES))
def test_base_sweep_matches_direct_solve(n, case):
    tau, omega, g, B, A, Bs, Bt, diffusion, F0, mu0 = TI._column(n, case)
    recs, mu, wt, tau_tot = TI.layer_records(n, tau, omega, g, B, F0, mu0, 2 * n)
    bottom = TI._bottom(n, recs, mu, wt, tau, B, A, Bs, diffusion, F0, mu0, tau_tot)
    ref = TI.direct_solve(n, recs, mu, wt, Bt * np.ones(n), bottom)
    got = _fluxes(n, wt, mu, *base_sweep(n, recs, Bt * np.ones(n), bottom))
    for name, a, b in zip(("flux_up", "flux_down", "mean_intensity"), got, ref):
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err <= TOL, (case, n, name, err)


def _deck_column(seed, second):
    """16-stream column with unstable starts at layer 5 (tau 0.3, omega 3e-4)
    and at layer `second` (tau 2.6, omega 1e-5); ordinary layers elsewhere."""
    n, L = 16, 10
    rng = np.random.default_rng(seed)
    tau = 10 ** rng.uniform(-3, -0.5, L)
    omega = rng.uniform(0.05, 0.9, L)
    tau[5], omega[5] = 0.3, 3e-4
    tau[second], omega[second] = 2.6, 1e-5
    g = rng.uniform(0.2, 0.85, L)
    B = np.linspace(1.0, 2.2, L + 1)
    A, Bs, Bt, F0, mu0 = 0.3, 1.1, 0.35, 1.0, 0.55
    recs, mu, wt, tau_tot = TI.layer_records(n, tau, omega, g, B, F0, mu0, 2 * n)
    bottom = TI._bottom(n, recs, mu, wt, tau, B, A, Bs, False, F0, mu0, tau_tot)
    ref = TI.direct_solve(n, recs, mu, wt, Bt * np.ones(n), bottom)
    base = _fluxes(n, wt, mu, *base_sweep(n, recs, Bt * np.ones(n), bottom))
    orc = TI._oracle_column(n, tau, omega, g, B, A, Bs, Bt, False, F0, mu0, 2 * n)
    e_base = max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(base[:2], ref[:2]))
    e_orc = max(np.abs(orc[k][0] - b).max() / np.abs(b).max()
                for k, b in zip(("flux_up", "flux_down"), ref[:2]))
    unstable = [l for l in range(L) if omega[l] < 0.01 and
                tau[l] * 2.0 ** -max(int(np.log2(tau[l])) + 4, 1) > oracle.gauss_legendre(n)[0][0]]
    assert unstable == sorted({5, second}), unstable
    return e_base, e_orc


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_consecutive_unstable_layers_need_the_composite_form(seed):
    e_base, e_orc = _deck_column(seed, 6)
    print(f"seed {seed}: base sweep {e_base:.1e}, reference adding {e_orc:.1e} from the direct solve")
    assert e_base > 10.0 * e_orc, (e_base, e_orc)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("second", [7, 8])
def test_separated_unstable_layers_keep_the_base_sweep(seed, second):
    e_base, e_orc = _deck_column(seed, second)
    assert e_base <= max(2.0 * e_orc, 1e-10), (e_base, e_orc)
