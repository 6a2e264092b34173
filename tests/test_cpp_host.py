"""The C++ host layer (include/adrt_hip.hpp) through its test driver
tests/cpp/adrt_cpp_check, against the CPU oracle.

* adrt::ADConfig::validate rejects what the reference rejects, with the
  reference's messages (src/adding_doubling.hpp:123-174) -- checked on the CPU,
  the driver throws before touching the GPU;
* adrt::solve (one column on the GPU, FP64) at the parity bar of the batched
  path: within max(1e-10, 4 x the FP64 oracle's own error) of the
  extended-precision oracle;
* the temperature Jacobians of adrt::solve (batched unit-source solves) against
  the oracle's forward-mode computeTemperatureJacobian (src/solver.cpp:117-341)
  to 1e-8 relative, for temperature-driven and Planck-level-driven columns;
* adrt::hip::solveBatch with temperatures on the device (band Planck in the
  library) in TOA mode and all-interfaces mode against the oracle per column;
* adrt::hip::solveBatchFromCoefficients (raw coefficients) in both modes
  against oracle.raw_optics + the oracle solve;
* adrt::hip::solveBatchHost (host vectors in, TOA results out) bit for bit
  against adrt::hip::solveBatch on the same inputs, and against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "adrt_cpp_check")
OUTS = ("flux_up", "flux_down", "mean_intensity", "flux_divergence", "flux_direct")
JACS = tuple(k + "_temperature_jac" for k in ("flux_up", "flux_down", "mean_intensity",
                                              "flux_divergence"))

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER),
                                reason="tests/cpp/adrt_cpp_check not built (run build())")


def _fmt(xs):
    return " ".join(repr(float(x)) for x in np.ravel(xs))


def run_driver(text):
    p = subprocess.run([DRIVER], input=text, capture_output=True, text=True, timeout=120)
    out = {}
    for line in p.stdout.splitlines():
        name, _, rest = line.partition(" ")
        if name in ("error", "failure"):
            out[name] = rest
        else:
            out[name] = np.array([float(v) for v in rest.split()])
    assert p.returncode == 0 and "failure" not in out, (p.returncode, p.stdout, p.stderr)
    return out


def column_case(L, N, tau, ssa, chi, thermal=False, delta_m=False, diffusion=False,
                from_bottom=False, jac=False, albedo=0.0, Bs=0.0, Ts=-1.0, Bt=0.0, Tt=-1.0,
                F0=0.0, mu0=1.0, lo=0.0, hi=0.0, T=(), B=()):
    chi = np.atleast_2d(np.asarray(chi, float))
    text = (f"column\n{L} {N} {int(thermal)} {int(delta_m)} {int(diffusion)} {int(from_bottom)} "
            f"{int(jac)}\n{albedo!r} {Bs!r} {Ts!r} {Bt!r} {Tt!r} {F0!r} {mu0!r} {lo!r} {hi!r}\n"
            f"{_fmt(tau)}\n{_fmt(ssa)}\n{chi.shape[1]} {_fmt(chi)}\n{len(T)} {_fmt(T)}\n"
            f"{len(B)} {_fmt(B)}\n")
    kw = dict(num_layers=L, num_quadrature=N, delta_tau=tau, single_scat_albedo=ssa,
              phase_function_moments=chi, use_thermal_emission=thermal, use_delta_m=delta_m,
              use_diffusion_lower_bc=diffusion, index_from_bottom=from_bottom,
              compute_temperature_jacobian=jac, surface_albedo=albedo, surface_emission=Bs,
              surface_temperature=Ts, top_emission=Bt, top_temperature=Tt, solar_flux=F0,
              solar_mu=mu0, wavenumber_low=lo, wavenumber_high=hi,
              temperature=np.asarray(T, float) if len(T) else None,
              planck_levels=np.asarray(B, float) if len(B) else None)
    return text, kw


INVALID = [
    dict(L=0, tau=[], ssa=[], chi=np.zeros((0, 4))),
    dict(N=1),
    dict(tau=[-0.1, 0.2, 0.3]),
    dict(ssa=[0.5, 1.2, 0.3]),
    dict(albedo=1.5),
    dict(F0=1.0, mu0=0.0),
    dict(thermal=True, T=[300.0, 280.0], lo=500.0, hi=600.0),
    dict(thermal=True, T=[300.0, -1.0, 250.0, 240.0], lo=500.0, hi=600.0),
    dict(thermal=True, T=[300.0, 280.0, 250.0, 240.0], lo=600.0, hi=500.0),
    dict(Ts=300.0),
    dict(Tt=200.0),
    dict(diffusion=True),
]


@pytest.mark.parametrize("i", range(len(INVALID)))
def test_validate_messages_match_reference(i):
    c = dict(L=3, N=4, tau=[0.1, 0.2, 0.3], ssa=[0.5, 0.9, 0.3], chi=[[1.0, 0.3, 0.1, 0.0]] * 3)
    c.update(INVALID[i])
    text, kw = column_case(**c)
    got = run_driver(text)
    with pytest.raises(oracle.OracleError) as e:
        oracle.solve(**kw)
    assert got.get("error") == str(e.value.args[1] if len(e.value.args) > 1 else e.value), \
        (got, e.value)


def _hg(g, M, L):
    return np.tile(g ** np.arange(M), (L, 1))


def _cases():
    rng = np.random.default_rng(5)
    L = 7
    tau = 10 ** rng.uniform(-2, 0.7, L)
    ssa = rng.uniform(0.2, 0.97, L)
    T = np.linspace(210.0, 300.0, L + 1)
    B = rng.uniform(0.5, 2.0, L + 1)
    return {
        "hg_solar_n16": dict(L=L, N=16, tau=tau, ssa=ssa, chi=_hg(0.7, 32, L), F0=1.0, mu0=0.6,
                             albedo=0.2),
        "thermal_T_n8": dict(L=L, N=8, tau=tau, ssa=ssa, chi=_hg(0.5, 16, L), thermal=True, T=T,
                             lo=600.0, hi=700.0, albedo=0.1, Ts=305.0, Tt=0.0, jac=True),
        "thermal_T_bottom_first": dict(L=L, N=4, tau=tau, ssa=ssa, chi=_hg(0.4, 8, L),
                                       thermal=True, T=T, lo=900.0, hi=950.0, albedo=0.3,
                                       from_bottom=True, jac=True),
        "planck_levels_delta_m": dict(L=L, N=8, tau=tau, ssa=ssa, chi=_hg(0.8, 17, L),
                                      delta_m=True, B=B, Bs=1.3, Bt=0.4, albedo=0.15, F0=0.5,
                                      mu0=0.4, jac=True),
        "diffusion_bc_n16": dict(L=L, N=16, tau=tau, ssa=ssa, chi=_hg(0.6, 32, L), thermal=True,
                                 T=T, lo=700.0, hi=720.0, diffusion=True, jac=True),
        "n32_solar": dict(L=L, N=32, tau=tau, ssa=ssa, chi=_hg(0.85, 64, L), F0=1.0, mu0=0.5,
                          albedo=0.25),
        "n32_thermal_jacobian": dict(L=L, N=32, tau=tau, ssa=ssa, chi=_hg(0.7, 64, L),
                                     thermal=True, T=T, lo=650.0, hi=700.0, albedo=0.2,
                                     Ts=300.0, F0=1.0, mu0=0.5, jac=True),
        "n6_padded_jacobian": dict(L=L, N=6, tau=tau, ssa=ssa, chi=_hg(0.5, 12, L),
                                   thermal=True, T=T, lo=800.0, hi=850.0, albedo=0.05,
                                   from_bottom=True, jac=True),
    }


CASES = _cases()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_column_solve_and_jacobians(name):
    text, kw = column_case(**CASES[name])
    got = run_driver(text)
    assert "error" not in got, got
    ref = oracle.solve(**kw)
    ext = oracle.solve(**kw, extended=True)
    for k in OUTS:
        g, r, e = got[k], ref[k], ext[k]
        scale = np.maximum(np.abs(e), 1e-300)
        err_g = np.abs(g - e) / scale
        err_o = np.abs(r - e) / scale
        bound = np.maximum(1e-10, 4.0 * err_o)
        if k == "flux_divergence":  # J ~ B cancels: scale by the terms
            bound = np.maximum(bound, 1e-10 * (np.abs(ext["mean_intensity"]) + 1.0) / scale)
        assert np.all(err_g <= bound), (name, k, float(err_g.max()))
    if CASES[name].get("jac"):
        for k in JACS:
            g = got[k].reshape(ref[k].shape)
            r = ref[k]
            tol = 1e-8 * np.abs(r).max() + 1e-8 * np.abs(r)
            assert np.all(np.abs(g - r) <= tol), (name, k, float(np.abs(g - r).max()),
                                                  float(np.abs(r).max()))


@pytest.mark.gpu
@pytest.mark.parametrize("all_interfaces", [False, True])
def test_batch_with_device_planck(all_interfaces):
    rng = np.random.default_rng(9)
    W, L, N, M = 5, 6, 8, 16
    tau = 10 ** rng.uniform(-2, 0.5, (W, L))
    ssa = rng.uniform(0.1, 0.95, (W, L))
    chi = 0.6 ** np.arange(M)
    mom = np.tile(chi, (L, 1))
    T = np.linspace(220.0, 295.0, L + 1)[None, :] + rng.normal(0, 3, (W, L + 1))
    lo, hi = 800.0, 850.0
    Ts = 300.0
    text = (f"batch\n{W} {L} {N} {M} 1 1 0 0 0 {int(all_interfaces)}\n"
            f"0.2 0.0 {Ts!r} 0.0 -1.0 1.0 0.7 {lo!r} {hi!r} 1.0\n"
            f"{_fmt(tau)}\n{_fmt(ssa)}\n{_fmt(mom)}\n{T.size} {_fmt(T)}\n0\n")
    got = run_driver(text)
    assert "error" not in got, got
    for w in range(W):
        ref = oracle.solve(L, N, tau[w], ssa[w], mom, temperature=T[w], use_thermal_emission=True,
                           surface_albedo=0.2, surface_temperature=Ts, solar_flux=1.0,
                           solar_mu=0.7, wavenumber_low=lo, wavenumber_high=hi)
        for k in ("flux_up", "flux_down", "flux_direct") + (
                ("mean_intensity", "flux_divergence") if all_interfaces else ()):
            g = got[k].reshape(W, -1)[w]
            r = ref[k] if all_interfaces else ref[k][[L if k == "flux_direct" else 0]]
            atol = 1e-13 * np.abs(r).max()
            if k == "flux_divergence":  # 4 pi (1 - omega) (J - B): J ~ B cancels
                Bl = np.array([oracle.planck(lo, hi, t) for t in T[w]])
                atol = 1e-10 * 4 * np.pi * (np.abs(ref["mean_intensity"]) + Bl)
            bad = np.abs(g - r) > 1e-10 * np.abs(r) + atol
            assert not bad.any(), (w, k, g[bad], r[bad])


@pytest.mark.gpu
@pytest.mark.parametrize("all_interfaces", [False, True])
def test_batch_from_coefficients(all_interfaces):
    rng = np.random.default_rng(13)
    W, L, N, M = 7, 5, 8, 16
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(1e5, 6e5, L))])
    # omega >= 0.02 everywhere: no unstable doubling start (the C-ABI raw path
    # is held to the oracle in that regime by tests/test_gpu_raw.py)
    a = 10 ** rng.uniform(-9, -5.5, (L + 1, W))
    s = a * rng.uniform(0.05, 1.0, (L + 1, W))
    T = np.linspace(295.0, 200.0, L + 1)
    nu = np.linspace(500.0, 1800.0, W)
    mom = 0.6 ** np.arange(M)[None, :].repeat(L, 0)
    cloud = np.zeros((L, W))
    cloud[1] = rng.uniform(0, 1e-3, W)
    Ts, Tt = 301.0, -1.0
    text = (f"raw\n{W} {L} {N} {M} 1 0 0 {int(all_interfaces)} 1\n"
            f"0.2 {Ts!r} {Tt!r} 1.0 0.6\n{_fmt(a)}\n{_fmt(s)}\n{_fmt(z)}\n{_fmt(T)}\n"
            f"{_fmt(nu)}\n{_fmt(mom)}\n{_fmt(cloud)}\n")
    got = run_driver(text)
    assert "error" not in got, got
    r = oracle.raw_optics(a, s, z, T, nu, cloud_optical_depth=cloud, surface_temperature=Ts,
                          top_temperature=Tt)
    ref = oracle.solve_batch(r["delta_tau"], r["single_scat_albedo"], mom, N,
                             planck_levels=r["planck_levels"],
                             surface_emission=r["surface_emission"],
                             top_emission=r["top_emission"], surface_albedo=0.2, solar_flux=1.0,
                             solar_mu=0.6)
    for k in ("flux_up", "flux_down", "flux_direct") + (
            ("mean_intensity", "flux_divergence") if all_interfaces else ()):
        g = got[k].reshape(W, -1)
        rk = ref[k] if all_interfaces else ref[k][:, [L if k == "flux_direct" else 0]]
        atol = 1e-13 * np.abs(rk).max()
        if k == "flux_divergence":  # 4 pi (1 - omega) (J - B): J ~ B cancels
            atol = 1e-10 * 4 * np.pi * (np.abs(ref["mean_intensity"]) + r["planck_levels"])
        bad = np.abs(g - rk) > 1e-10 * np.abs(rk) + atol
        assert not bad.any(), (k, g[bad], rk[bad])


@pytest.mark.gpu
@pytest.mark.parametrize("thermal", [True, False])
def test_batch_host_matches_device_batch(thermal):
    """solveBatchHost (src_cuda/cuda_solver.cuh:91-98) uploads, solves and
    downloads: the same numbers as solveBatch on device buffers, bit for bit."""
    rng = np.random.default_rng(17)
    W, L, N, M = 6, 5, 8, 16
    tau = 10 ** rng.uniform(-2, 0.5, (W, L))
    ssa = rng.uniform(0.1, 0.95, (W, L))
    mom = rng.uniform(0.3, 0.8, (W, L))[..., None] ** np.arange(M)
    T = np.linspace(230.0, 290.0, L + 1)[None, :] + rng.normal(0, 2, (W, L + 1))  # [nwav][nlev]
    lo, hi = 700.0, 760.0
    head = (f"{W} {L} {N} {M} 0 {int(thermal)} 0 0 0 0\n"
            f"0.25 0.0 {295.0 if thermal else -1.0!r} 0.0 -1.0 1.3 0.65 {lo!r} {hi!r} 1.0\n"
            f"{_fmt(tau)}\n{_fmt(ssa)}\n{_fmt(mom)}\n"
            f"{T.size if thermal else 0} {_fmt(T) if thermal else ''}\n0\n")
    dev = run_driver("batch\n" + head)
    host = run_driver("host\n" + head)
    assert "error" not in dev and "error" not in host, (dev, host)
    for k in ("flux_up", "flux_down", "flux_direct"):
        assert np.array_equal(dev[k], host[k]), k
    for w in range(W):
        kw = dict(temperature=T[w], use_thermal_emission=True, surface_temperature=295.0,
                  wavenumber_low=lo, wavenumber_high=hi) if thermal else {}
        ref = oracle.solve(L, N, tau[w], ssa[w], mom[w], surface_albedo=0.25, solar_flux=1.3,
                           solar_mu=0.65, **kw)
        for k, lev in (("flux_up", 0), ("flux_down", 0), ("flux_direct", L)):
            r = ref[k][lev]
            assert abs(host[k][w] - r) <= 1e-10 * abs(r) + 1e-13, (w, k, host[k][w], r)


def test_batch_host_rejects_short_vectors():
    """A host vector shorter than its device extent is refused before any copy
    (checked on the host; no GPU work is started)."""
    W, L, N, M = 2, 3, 4, 8
    text = (f"host\n{W} {L} {N} {M} 1 1 0 0 0 0\n"
            f"0.1 0.0 290.0 0.0 -1.0 0.0 1.0 700.0 710.0 1.0\n"
            f"{_fmt(np.full(W * L, 0.3))}\n{_fmt(np.full(W * L, 0.5))}\n"
            f"{_fmt(np.ones(L * M))}\n{L + 1} {_fmt(np.full(L + 1, 250.0))}\n0\n")
    got = run_driver(text)
    assert "temperature has 4 values, expected 8" in got.get("error", ""), got
