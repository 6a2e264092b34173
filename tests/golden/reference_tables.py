"""Published benchmark values held by the reference's own test suite.

Data only: each case records a column configuration and the expected fluxes
that the reference test file ``tests/test_ad_solver.cpp`` checks (CDISORT /
Van de Hulst / Sweigart / Garcia-Siewert / Ozisik-Shouman tables), with the
tolerance that test uses.  ``phase`` names how the moments are filled, with
the semantics of ``ADConfig::allocate`` + ``set*`` (reference
``src/adding_doubling.hpp:107-246``); ``moments(...)`` below expands it.

Each check is ``(output, interface, expected, abs_tol)``.
"""
import math

PI = math.pi

# Haze-L Garcia-Siewert moments (2l+1)chi_l, l = 1..82 (test_ad_solver.cpp:28-41)
HAZE_L = [
    2.41260, 3.23047, 3.37296, 3.23150, 2.89350, 2.49594, 2.11361, 1.74812,
    1.44692, 1.17714, 0.96643, 0.78237, 0.64114, 0.51966, 0.42563, 0.34688,
    0.28351, 0.23317, 0.18963, 0.15788, 0.12739, 0.10762, 0.08597, 0.07381,
    0.05828, 0.05089, 0.03971, 0.03524, 0.02720, 0.02451, 0.01874, 0.01711,
    0.01298, 0.01198, 0.00904, 0.00841, 0.00634, 0.00592, 0.00446, 0.00418,
    0.00316, 0.00296, 0.00225, 0.00210, 0.00160, 0.00150, 0.00115, 0.00107,
    0.00082, 0.00077, 0.00059, 0.00055, 0.00043, 0.00040, 0.00031, 0.00029,
    0.00023, 0.00021, 0.00017, 0.00015, 0.00012, 0.00011, 0.00009, 0.00008,
    0.00006, 0.00006, 0.00005, 0.00004, 0.00004, 0.00003, 0.00003, 0.00002,
    0.00002, 0.00002, 0.00001, 0.00001, 0.00001, 0.00001, 0.00001, 0.00001,
    0.00001, 0.00001,
]

# DGIS moments (test_ad_solver.cpp:788-798)
DGIS = [1.0, 2.00916 / 3.0, 1.56339 / 5.0, 0.67407 / 7.0, 0.22215 / 9.0,
        0.04725 / 11.0, 0.00671 / 13.0, 0.00068 / 15.0, 0.00005 / 17.0]


def moments(kind, nmom, g=0.0, f=1.0, g2=0.0):
    """Moment vector of length ``nmom`` for a named phase function."""
    chi = [0.0] * nmom
    chi[0] = 1.0
    if kind == "isotropic":
        pass
    elif kind == "rayleigh":
        if nmom > 2:
            chi[2] = 0.1
    elif kind == "hg":
        gk = 1.0
        for k in range(nmom):
            chi[k] = gk
            gk *= g
    elif kind == "dhg":
        a = b = 1.0
        for k in range(nmom):
            chi[k] = f * a + (1.0 - f) * b
            a *= g
            b *= g2
    elif kind == "haze":
        for k in range(1, min(82, nmom - 1) + 1):
            chi[k] = HAZE_L[k - 1] / (2 * k + 1)
    elif kind == "dgis":
        for k in range(min(9, nmom)):
            chi[k] = DGIS[k]
    else:
        raise ValueError(kind)
    return chi


def _t(x, rel):
    return rel * abs(x)


# Each case: name, config (oracle.solve keyword arguments except moments),
# layer phase specs, and checks.
CASES = [
    dict(name="Isotropic1a", line=261, nq=8, layers=[(0.03125, 0.2, ("isotropic",))],
         cfg=dict(solar_flux=10 * PI, solar_mu=0.1),
         checks=[("flux_direct", 0, 3.1416, _t(3.1416, 2e-2)),
                 ("flux_direct", 1, 2.2984, _t(2.2984, 2e-2)),
                 ("flux_down", 0, 0.0, 1e-4),
                 ("flux_up", 0, 7.9945e-02, _t(7.9945e-02, 2e-2)),
                 ("flux_down", 1, 7.9411e-02, _t(7.9411e-02, 2e-2)),
                 ("flux_up", 1, 0.0, 1e-4)]),
    dict(name="Isotropic1b", line=283, nq=8, layers=[(0.03125, 1.0, ("isotropic",))],
         cfg=dict(solar_flux=10 * PI, solar_mu=0.1),
         checks=[("flux_up", 0, 0.42292, _t(0.42292, 2e-2)),
                 ("flux_down", 1, 0.42023, _t(0.42023, 2e-2))]),
    dict(name="Rayleigh2a", line=352, nq=8, layers=[(0.2, 0.5, ("rayleigh",))],
         cfg=dict(solar_flux=PI, solar_mu=0.080442),
         checks=[("flux_direct", 0, 2.52716e-01, _t(2.52716e-01, 2e-2)),
                 ("flux_direct", 1, 2.10311e-02, _t(2.10311e-02, 2e-2)),
                 ("flux_down", 0, 0.0, 1e-6), ("flux_up", 1, 0.0, 1e-6),
                 ("flux_up", 0, 5.35063e-02, _t(5.35063e-02, 2e-2)),
                 ("flux_down", 1, 4.41794e-02, _t(4.41794e-02, 2e-2))]),
    dict(name="Rayleigh2b", line=374, nq=8, layers=[(0.2, 1.0, ("rayleigh",))],
         cfg=dict(solar_flux=PI, solar_mu=0.080442),
         checks=[("flux_up", 0, 1.25561e-01, _t(1.25561e-01, 2e-2)),
                 ("flux_down", 1, 1.06123e-01, _t(1.06123e-01, 2e-2))]),
    dict(name="Rayleigh2c", line=394, nq=8, layers=[(5.0, 0.5, ("rayleigh",))],
         cfg=dict(solar_flux=PI, solar_mu=0.080442),
         checks=[("flux_up", 0, 6.24730e-02, _t(6.24730e-02, 2e-2)),
                 ("flux_down", 1, 2.51683e-04, 3e-5)]),
    dict(name="Rayleigh2d", line=414, nq=8, layers=[(5.0, 1.0, ("rayleigh",))],
         cfg=dict(solar_flux=PI, solar_mu=0.080442),
         checks=[("flux_up", 0, 2.25915e-01, _t(2.25915e-01, 2e-2)),
                 ("flux_down", 1, 2.68008e-02, _t(2.68008e-02, 2e-2))]),
    dict(name="HenyeyGreenstein3a", line=440, nq=8, layers=[(1.0, 1.0, ("hg", 0.75))],
         cfg=dict(solar_flux=PI, solar_mu=1.0),
         checks=[("flux_direct", 1, 1.15573, _t(1.15573, 2e-2)),
                 ("flux_up", 0, 2.47374e-01, _t(2.47374e-01, 2e-2)),
                 ("flux_down", 1, 1.73849, _t(1.73849, 2e-2)),
                 ("flux_up", 1, 0.0, 1e-6)]),
    dict(name="HenyeyGreenstein3b", line=461, nq=8, layers=[(8.0, 1.0, ("hg", 0.75))],
         cfg=dict(solar_flux=PI, solar_mu=1.0),
         checks=[("flux_direct", 1, 1.05389e-03, _t(1.05389e-03, 2e-2)),
                 ("flux_up", 0, 1.59096, _t(1.59096, 5e-2)),
                 ("flux_down", 1, 1.54958, _t(1.54958, 5e-2))]),
    dict(name="HazeGarciaSiewert4a", line=487, nq=16, layers=[(1.0, 1.0, ("haze",))],
         cfg=dict(solar_flux=PI, solar_mu=1.0),
         checks=[("flux_up", 0, 1.73223e-01, _t(1.73223e-01, 2e-2)),
                 ("flux_down", 1, 1.81264, _t(1.81264, 2e-2))]),
    dict(name="HazeGarciaSiewert4b", line=507, nq=16, layers=[(1.0, 0.9, ("haze",))],
         cfg=dict(solar_flux=PI, solar_mu=1.0),
         checks=[("flux_up", 0, 1.23665e-01, _t(1.23665e-01, 2e-2)),
                 ("flux_down", 1, 1.51554, _t(1.51554, 2e-2))]),
    dict(name="HazeGarciaSiewert4c", line=527, nq=16, layers=[(1.0, 0.9, ("haze",))],
         cfg=dict(solar_flux=PI, solar_mu=0.5),
         checks=[("flux_direct", 0, 1.57080, _t(1.57080, 2e-2)),
                 ("flux_direct", 1, 2.12584e-01, _t(2.12584e-01, 2e-2)),
                 ("flux_up", 0, 2.25487e-01, _t(2.25487e-01, 2e-2)),
                 ("flux_down", 1, 8.03294e-01, _t(8.03294e-01, 2e-2))]),
    dict(name="MultiLayer1_TwoLayerRayleigh", line=553, nq=8,
         layers=[(0.1, 0.5, ("rayleigh",)), (0.2, 0.8, ("rayleigh",))],
         cfg=dict(solar_flux=1.0, solar_mu=0.5),
         checks=[("flux_up", 0, 6.593572e-02, _t(6.593572e-02, 2e-2)),
                 ("flux_direct", 1, 4.093654e-01, _t(4.093654e-01, 2e-2)),
                 ("flux_down", 1, 2.327019e-02, _t(2.327019e-02, 2e-2)),
                 ("flux_up", 1, 5.456810e-02, _t(5.456810e-02, 2e-2)),
                 ("flux_direct", 2, 2.744058e-01, _t(2.744058e-01, 2e-2)),
                 ("flux_down", 2, 6.715049e-02, _t(6.715049e-02, 2e-2)),
                 ("flux_up", 2, 0.0, 1e-6)]),
    dict(name="MultiLayer2_TwoLayerHG", line=591, nq=8,
         layers=[(0.5, 1.0, ("hg", 0.75)), (0.5, 1.0, ("hg", 0.75))],
         cfg=dict(solar_flux=PI, solar_mu=1.0),
         checks=[("flux_up", 0, 2.473744e-01, _t(2.473744e-01, 2e-2)),
                 ("flux_direct", 1, 1.905472e+00, _t(1.905472e+00, 2e-2)),
                 ("flux_down", 1, 1.149116e+00, _t(1.149116e+00, 5e-2)),
                 ("flux_up", 1, 1.603704e-01, _t(1.603704e-01, 5e-2)),
                 ("flux_direct", 2, 1.155727e+00, _t(1.155727e+00, 2e-2)),
                 ("flux_down", 2, 1.738491e+00, _t(1.738491e+00, 2e-2))]),
    dict(name="MultiLayer3_ThreeLayerIsotropic", line=629, nq=8,
         layers=[(0.5, 0.5, ("isotropic",)), (1.0, 0.9, ("isotropic",)),
                 (0.5, 0.5, ("isotropic",))],
         cfg=dict(solar_flux=1.0, solar_mu=0.5, surface_albedo=0.1),
         checks=[("flux_up", 0, 1.133613e-01, _t(1.133613e-01, 2e-2)),
                 ("flux_direct", 1, 1.839397e-01, _t(1.839397e-01, 2e-2)),
                 ("flux_down", 1, 6.416506e-02, _t(6.416506e-02, 2e-2)),
                 ("flux_up", 1, 1.031752e-01, _t(1.031752e-01, 2e-2)),
                 ("flux_direct", 2, 2.489353e-02, _t(2.489353e-02, 2e-2)),
                 ("flux_down", 2, 8.605320e-02, _t(8.605320e-02, 2e-2)),
                 ("flux_up", 2, 1.536433e-02, _t(1.536433e-02, 2e-2)),
                 ("flux_direct", 3, 9.157819e-03, _t(9.157819e-03, 2e-2)),
                 ("flux_down", 3, 4.944670e-02, _t(4.944670e-02, 2e-2)),
                 ("flux_up", 3, 5.860452e-03, _t(5.860452e-03, 2e-2))]),
    dict(name="DiffuseIllumination8a", line=672, nq=4,
         layers=[(0.25, 0.5, ("isotropic",)), (0.25, 0.3, ("isotropic",))],
         cfg=dict(top_emission=1.0 / PI),
         checks=[("flux_down", 0, 1.0, 2e-2),
                 ("flux_up", 0, 9.29633e-02, _t(9.29633e-02, 2e-2)),
                 ("flux_down", 1, 7.22235e-01, _t(7.22235e-01, 2e-2)),
                 ("flux_up", 1, 2.78952e-02, _t(2.78952e-02, 2e-2)),
                 ("flux_down", 2, 5.13132e-01, _t(5.13132e-01, 2e-2)),
                 ("flux_up", 2, 0.0, 1e-6)]),
    dict(name="DiffuseIllumination8b", line=702, nq=4,
         layers=[(0.25, 0.8, ("isotropic",)), (0.25, 0.95, ("isotropic",))],
         cfg=dict(top_emission=1.0 / PI),
         checks=[("flux_up", 0, 2.25136e-01, _t(2.25136e-01, 2e-2)),
                 ("flux_down", 1, 7.95332e-01, _t(7.95332e-01, 2e-2)),
                 ("flux_up", 1, 1.26349e-01, _t(1.26349e-01, 2e-2)),
                 ("flux_down", 2, 6.50417e-01, _t(6.50417e-01, 2e-2))]),
    dict(name="DiffuseIllumination8c", line=726, nq=4,
         layers=[(1.0, 0.8, ("isotropic",)), (2.0, 0.95, ("isotropic",))],
         cfg=dict(top_emission=1.0 / PI),
         checks=[("flux_up", 0, 3.78578e-01, _t(3.78578e-01, 2e-2)),
                 ("flux_down", 1, 4.86157e-01, _t(4.86157e-01, 2e-2)),
                 ("flux_up", 1, 2.43397e-01, _t(2.43397e-01, 2e-2)),
                 ("flux_down", 2, 1.59984e-01, _t(1.59984e-01, 2e-2))]),
    dict(name="SixLayer9a_Isotropic", line=756, nq=4,
         layers=[(float(lc + 1), 0.65 + lc * 0.05, ("isotropic",)) for lc in range(6)],
         cfg=dict(top_emission=1.0 / PI),
         checks=[("flux_down", 0, 1.0, 2e-2),
                 ("flux_up", 0, 2.279734e-01, _t(2.279734e-01, 5e-2)),
                 ("flux_up", 1, 8.750978e-02, _t(8.750978e-02, 5e-2)),
                 ("flux_up", 6, 0.0, 1e-6)]),
    dict(name="SixLayer9b_Anisotropic", line=786, nq=4,
         layers=[(float(lc + 1), 0.65 + lc * 0.05, ("dgis",)) for lc in range(6)],
         cfg=dict(top_emission=1.0 / PI),
         checks=[("flux_up", 0, 1.000789e-01, _t(1.000789e-01, 2e-2)),
                 ("flux_up", 6, 0.0, 1e-6)]),
    dict(name="CombinedSources11a", line=830, nq=8, layers=[(1.0, 0.9, ("isotropic",))],
         cfg=dict(solar_flux=1.0, solar_mu=0.5, top_emission=0.5 / PI, surface_albedo=0.5),
         checks=[("flux_direct", 0, 0.5, 1e-2), ("flux_down", 0, 0.5, 1e-2),
                 ("flux_up", 0, 5.013753e-01, _t(5.013753e-01, 2e-2)),
                 ("flux_direct", 1, 6.766764e-02, _t(6.766764e-02, 2e-2)),
                 ("flux_down", 1, 4.723631e-01, _t(4.723631e-01, 2e-2)),
                 ("flux_up", 1, 2.700154e-01, _t(2.700154e-01, 2e-2))]),
    dict(name="CombinedSources11b_LayerSplit", line=859, nq=8,
         layers=[(0.05, 0.9, ("isotropic",)), (0.45, 0.9, ("isotropic",)),
                 (0.50, 0.9, ("isotropic",))],
         cfg=dict(solar_flux=1.0, solar_mu=0.5, top_emission=0.5 / PI, surface_albedo=0.5),
         checks=[("flux_up", 0, 5.013753e-01, _t(5.013753e-01, 2e-2)),
                 ("flux_direct", 3, 6.766764e-02, _t(6.766764e-02, 2e-2)),
                 ("flux_down", 3, 4.723631e-01, _t(4.723631e-01, 2e-2)),
                 ("flux_up", 3, 2.700154e-01, _t(2.700154e-01, 2e-2))]),
]

# Thermal table cases need the Planck function for their level values; they are
# expressed as (lo, hi, T) triples expanded by the test with the oracle's Planck.
THERMAL_CASES = [
    dict(name="Thermal7a", line=894, nq=8, layers=[(1.0, 0.1, ("hg", 0.05))],
         planck=[(300.0, 800.0, 200.0), (300.0, 800.0, 300.0)],
         top=None, surface=None, cfg=dict(),
         checks=[("flux_up", 0, 8.62935618e+01, _t(8.62935618e+01, 2e-2)),
                 ("flux_down", 1, 1.21203517e+02, _t(1.21203517e+02, 2e-2))]),
    dict(name="Thermal7c", line=923, nq=6, layers=[(1.0, 0.5, ("hg", 0.8))],
         planck=[(0.01, 50000.0, 300.0), (0.01, 50000.0, 200.0)],
         top=(100.0, (0.01, 50000.0, 100.0)), surface=(0.01, 50000.0, 320.0),
         cfg=dict(solar_flux=200.0, solar_mu=0.5),
         checks=[("flux_direct", 0, 1.0e+02, _t(1.0e+02, 2e-2)),
                 ("flux_up", 0, 4.29571753e+02, _t(4.29571753e+02, 2e-2))]),
]
