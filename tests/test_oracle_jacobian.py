"""Oracle temperature Jacobian vs central finite differences of the oracle.

Protocol and case matrix of the reference's tests/test_temperature_jacobian.cpp
(h = 1e-3 T, rtol 3e-3, atol 1e-6 x flux scale, every level DOF plus the
surface DOF; folded surface column identically zero).
"""
import numpy as np
import pytest

import oracle
from golden.reference_tables import moments

OUTS = ("flux_up", "flux_down", "mean_intensity", "flux_divergence")


def base_thermal(nlay, nq, ssa, g=None, **kw):
    nm = 2 * nq + (1 if kw.get("use_delta_m") else 0)
    mom = [moments("hg", nm, g=g) if g is not None else moments("isotropic", nm)] * nlay
    c = dict(num_layers=nlay, num_quadrature=nq, use_thermal_emission=True,
             wavenumber_low=500.0, wavenumber_high=1500.0,
             delta_tau=[0.3 + 0.1 * l for l in range(nlay)], single_scat_albedo=[ssa] * nlay,
             temperature=[200.0 + 8.0 * l for l in range(nlay + 1)],
             phase_function_moments=mom)
    c.update(kw)
    return c


def run_case(cfg):
    ana = oracle.solve(**dict(cfg, compute_temperature_jacobian=True))
    nlay = cfg["num_layers"]
    scale = max(1.0, float(np.max(np.abs(ana["flux_up"]) + np.abs(ana["flux_down"]))))
    cols = list(range(nlay + 1))
    surf = cfg.get("surface_temperature", -1.0) >= 0.0
    if surf:
        cols.append(nlay + 1)
    else:
        assert np.all(np.abs(ana["flux_up_temperature_jac"][:, nlay + 1]) < 1e-12)
    for m in cols:
        c = dict(cfg)
        if m == nlay + 1:
            T0 = c["surface_temperature"]
        else:
            c["temperature"] = list(cfg["temperature"])
            T0 = c["temperature"][m]
        h = 1e-3 * max(abs(T0), 1.0)
        res = []
        for s in (+1, -1):
            if m == nlay + 1:
                c["surface_temperature"] = T0 + s * h
            else:
                c["temperature"][m] = T0 + s * h
            res.append(oracle.solve(**c))
        for k in OUTS:
            fd = (res[0][k] - res[1][k]) / (2 * h)
            a = ana[k + "_temperature_jac"][:, m]
            assert np.all(np.abs(a - fd) <= 1e-6 * scale + 3e-3 * np.abs(fd)), (k, m)


CASES = {
    "AbsorptionWithSurface": base_thermal(4, 8, 0.0),
    "ScatteringDistinctSurface": base_thermal(5, 8, 0.5, g=0.5, surface_albedo=0.3,
                                              surface_temperature=250.0),
    "HighScatteringDiffusionBC": base_thermal(5, 8, 0.9, g=0.7, use_diffusion_lower_bc=True),
    "WithSolarBeamInvariance": base_thermal(4, 8, 0.6, g=0.4, surface_albedo=0.2,
                                            surface_temperature=240.0, solar_flux=100.0,
                                            solar_mu=0.6),
    "QuadratureOrder4": base_thermal(4, 4, 0.7, g=0.5, surface_albedo=0.25,
                                     surface_temperature=245.0),
    "QuadratureOrder16": base_thermal(4, 16, 0.7, g=0.5, surface_albedo=0.25,
                                      surface_temperature=245.0),
    "DynamicPath6": base_thermal(5, 6, 0.6, g=0.5, surface_albedo=0.3, surface_temperature=250.0),
    "DynamicPath12": base_thermal(5, 12, 0.6, g=0.5, surface_albedo=0.3,
                                  surface_temperature=250.0),
    "DynamicDiffusion": base_thermal(5, 6, 0.9, g=0.7, use_diffusion_lower_bc=True),
    "ColdSpaceTop": base_thermal(4, 8, 0.6, g=0.5, surface_albedo=0.2, surface_temperature=250.0,
                                 top_temperature=0.0),
    "ExplicitTopAbsorbing": base_thermal(4, 8, 0.0, top_temperature=180.0),
    "ExplicitTopDynamic": base_thermal(5, 6, 0.6, g=0.5, surface_albedo=0.3,
                                       surface_temperature=250.0, top_temperature=210.0),
}


def _delta_m_case():
    c = base_thermal(5, 8, 0.8, g=0.8, use_delta_m=True, surface_albedo=0.1,
                     surface_temperature=255.0)
    c["delta_tau"] = [0.5] * 5
    c["temperature"] = [210.0 + 6.0 * l for l in range(6)]
    return c


CASES["DeltaMScaling"] = _delta_m_case()


@pytest.mark.parametrize("name", sorted(CASES))
def test_temperature_jacobian_fd(name):
    run_case(CASES[name])


def test_jacobian_index_from_bottom():
    """Reversed indexing reverses interface rows and level columns."""
    c = CASES["ScatteringDistinctSurface"]
    a = oracle.solve(**dict(c, compute_temperature_jacobian=True))
    r = dict(c, index_from_bottom=True, compute_temperature_jacobian=True,
             delta_tau=c["delta_tau"][::-1], temperature=c["temperature"][::-1])
    b = oracle.solve(**r)
    nlay = c["num_layers"]
    Ja = a["flux_up_temperature_jac"]
    Jb = b["flux_up_temperature_jac"]
    assert np.allclose(Ja[:, :nlay + 1], Jb[::-1, nlay::-1], rtol=1e-10, atol=1e-14)
    assert np.allclose(Ja[:, nlay + 1], Jb[::-1, nlay + 1], rtol=1e-10, atol=1e-14)
