"""CPU checks of the product boundary: the HIP library loads and exports every
symbol include/adrt_hip.h declares, its host utilities agree with the oracle,
and the synthetic workloads have the configured shapes.  No GPU call is made.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from addingdoublingrt_amd import _lib, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "adrt_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(adrt_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_status_strings():
    L = _lib.lib()
    assert L.adrt_version().decode().startswith("adrt")
    assert "num_layers" in L.adrt_status_string(1).decode()


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16, 32])
def test_quadrature_identical_to_oracle(n):
    from addingdoublingrt_amd import gauss_legendre
    mu, wt = gauss_legendre(n)
    om, ow = oracle.gauss_legendre(n)
    assert np.array_equal(mu, om) and np.array_equal(wt, ow)


@pytest.mark.parametrize("lo,hi,T", [(500, 600, 300), (0.01, 50000, 250), (1000, 1000.1, 260),
                                     (700, 700, 200), (500, 600, 0.0), (10, 30, 310)])
def test_host_planck_matches_oracle(lo, hi, T):
    from addingdoublingrt_amd import planck, planck_deriv
    assert planck(lo, hi, T) == pytest.approx(oracle.planck(lo, hi, T), rel=1e-15, abs=0)
    assert planck_deriv(lo, hi, T) == pytest.approx(oracle.planck_deriv(lo, hi, T), rel=1e-14,
                                                    abs=0)


def test_planck_rejects_bad_band():
    from addingdoublingrt_amd import AdrtError, planck
    with pytest.raises(AdrtError) as e:
        planck(600, 500, 300)
    assert e.value.code == 16


def test_workspace_rejects_bad_config_without_gpu():
    L = _lib.lib()
    cfg = _lib.BatchConfig(num_wavenumbers=4, num_layers=0, num_quadrature=8, num_moments=16)
    h = C.c_void_p()
    assert L.adrt_workspace_create(C.byref(cfg), 0, C.byref(h)) == 1
    cfg.num_layers = 3
    cfg.num_quadrature = 1
    assert L.adrt_workspace_create(C.byref(cfg), 0, C.byref(h)) == 2
    cfg.num_quadrature = 40
    assert L.adrt_workspace_create(C.byref(cfg), 0, C.byref(h)) == 17
    cfg.num_quadrature = 8
    cfg.surface_albedo = 1.5
    assert L.adrt_workspace_create(C.byref(cfg), 0, C.byref(h)) == 8
    cfg.surface_albedo = 0.1
    cfg.solar_flux = 1.0
    cfg.solar_mu = 0.0
    assert L.adrt_workspace_create(C.byref(cfg), 0, C.byref(h)) == 9


def test_raw_entry_rejects_bad_arguments_without_gpu():
    L = _lib.lib()
    cfg = _lib.BatchConfig(num_wavenumbers=4, num_layers=3, num_quadrature=8, num_moments=16,
                           solar_mu=1.0, spectrum_scaling=1.0)
    coef = _lib.DeviceCoefficients()
    out = _lib.DeviceOutputs()
    null = C.c_void_p()
    f = L.adrt_solve_batch_from_coefficients
    assert f(None, C.byref(coef), -1.0, -1.0, C.byref(out), 0, null, null) == 23
    assert f(C.byref(cfg), None, -1.0, -1.0, C.byref(out), 0, null, null) == 23
    assert f(C.byref(cfg), C.byref(coef), -1.0, -1.0, C.byref(out), 0, null, null) == 23
    cfg.num_quadrature = 1
    assert f(C.byref(cfg), C.byref(coef), -1.0, -1.0, C.byref(out), 0, null, null) == 2
    cfg.num_quadrature = 8
    cfg.num_wavenumbers = 0  # nothing to do: no pointer is touched
    assert f(C.byref(cfg), C.byref(coef), -1.0, -1.0, C.byref(out), 0, null, null) == 0


def test_raw_optics_entries_reject_bad_arguments_without_gpu():
    L = _lib.lib()
    coef = _lib.DeviceCoefficients()
    null = C.c_void_p()
    outs = [null] * 5
    f = L.adrt_raw_optics
    assert f(4, 3, None, -1.0, -1.0, *outs, null) == 23
    assert f(-1, 3, C.byref(coef), -1.0, -1.0, *outs, null) == 18
    assert f(4, 0, C.byref(coef), -1.0, -1.0, *outs, null) == 1
    assert f(4, 3, C.byref(coef), -1.0, -1.0, *outs, null) == 23
    assert f(0, 3, C.byref(coef), -1.0, -1.0, *outs, null) == 0  # empty: nothing launched
    g = L.adrt_raw_optics_adjoint
    cot, grad = _lib.RawOpticsCotangents(), _lib.RawGradients()
    assert g(4, 3, None, -1.0, -1.0, C.byref(cot), C.byref(grad), null) == 23
    assert g(4, 3, C.byref(coef), -1.0, -1.0, None, C.byref(grad), null) == 23
    assert g(4, 3, C.byref(coef), -1.0, -1.0, C.byref(cot), None, null) == 23
    assert g(-2, 3, C.byref(coef), -1.0, -1.0, C.byref(cot), C.byref(grad), null) == 18
    assert g(4, 0, C.byref(coef), -1.0, -1.0, C.byref(cot), C.byref(grad), null) == 1
    assert g(4, 3, C.byref(coef), -1.0, -1.0, C.byref(cot), C.byref(grad), null) == 23


@pytest.mark.parametrize("k,W,L,N", [(1, 1000, 50, 8), (2, 16384, 100, 16), (3, 8192, 80, 32),
                                     (4, 4096, 60, 8), (5, 65536, 100, 16)])
def test_workload_shapes(k, W, L, N):
    c = workloads.make(k, W=7)
    assert c.delta_tau.shape == (7, L) and c.N == N
    full = workloads.make(k) if W <= 4096 else None
    if full is not None:
        assert full.W == W
    assert np.all(c.delta_tau > 0) and np.all((c.single_scat_albedo >= 0) &
                                              (c.single_scat_albedo <= 1))
    assert c.phase_moments.shape[-1] >= 2 * N or k == 1


def test_workload_is_deterministic():
    a = workloads.make(2, W=5)
    b = workloads.make(2, W=5)
    assert np.array_equal(a.delta_tau, b.delta_tau) and np.array_equal(a.phase_moments,
                                                                       b.phase_moments)


def test_raw_optics_rejects_host_tensors():
    torch = pytest.importorskip("torch")
    from addingdoublingrt_amd import raw_optics

    L, W = 3, 2
    t = lambda *shape: torch.ones(shape, dtype=torch.float64)
    with pytest.raises(TypeError, match="device tensor"):
        raw_optics(t(L + 1, W), t(L + 1, W), t(L + 1), t(L + 1), t(W))
