"""Gradients of the raw-coefficient path (adrt_raw_optics_adjoint, chained with
the solve's reverse pass), through the C-ABI.

* ``raw_optics`` alone: torch.autograd.gradcheck (central differences of the
  HIP forward) over every input element, with clouds and boundary
  temperatures, with the "use the bottom / top level" defaults, and with a
  layer whose optical depth is clamped at zero.
* The forward of ``raw_optics`` equals oracle.raw_optics (the restatement of
  src_cuda/cuda_solver.cu:1131-1182) to a few ulp, and
  ``solve_from_coefficients`` agrees with ``solve_batch_from_coefficients``.
* ``solve_from_coefficients`` gradients against Richardson-extrapolated
  central differences of the oracle (oracle.raw_optics feeding the
  extended-precision oracle.solve_batch), to 1e-7 relative, for single entries of
  the coefficients, the cloud depths and the temperatures, at N = 8 and 16;
  at N = 32 for the temperatures (the only inputs with gradients there).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from addingdoublingrt_amd import (BatchConfig, raw_optics, solve_batch_from_coefficients,  # noqa: E402
                                  solve_from_coefficients)

pytestmark = pytest.mark.gpu
OUTS = ("flux_up", "flux_down", "mean_intensity", "flux_divergence")


def _t(x, grad=False):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda",
                        requires_grad=grad)


def _order_one(W, L, seed, cloud):
    """Coefficients and depths of order one (so that gradcheck's absolute step
    is small against every input)."""
    rng = np.random.default_rng(seed)
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.2, 1.0, L))])
    a = rng.uniform(0.1, 1.0, (L + 1, W))
    s = rng.uniform(0.1, 1.0, (L + 1, W))
    T = rng.uniform(200.0, 320.0, L + 1)
    nu = rng.uniform(300.0, 2500.0, W)
    cl = rng.uniform(0.0, 0.5, (L, W)) if cloud else None
    return z, a, s, T, nu, cl


@pytest.mark.parametrize("cloud,Ts,Tt", [(True, 295.0, 180.0), (False, -1.0, -1.0),
                                         (True, -1.0, 210.0)])
def test_raw_optics_gradcheck(cloud, Ts, Tt):
    print(f"gradcheck cloud {cloud} Ts {Ts} Tt {Tt}", flush=True)
    W, L = 3, 4
    z, a, s, T, nu, cl = _order_one(W, L, 7, cloud)
    args = [_t(a, True), _t(s, True), _t(z), _t(T, True), _t(nu),
            _t(cl, True) if cloud else None, _t(Ts, True), _t(Tt, True)]

    def f(a_, s_, T_, cl_, ts_, tt_):
        r = raw_optics(a_, s_, args[2], T_, args[4], cl_, ts_, tt_)
        return tuple(r.values())

    inputs = (args[0], args[1], args[3], args[5], args[6], args[7])
    assert torch.autograd.gradcheck(f, inputs, eps=1e-6, atol=1e-9, rtol=1e-7)


@pytest.mark.parametrize("cloud,Ts,Tt", [(True, 295.0, -1.0), (False, -1.0, 180.0)])
def test_raw_optics_matches_oracle(cloud, Ts, Tt):
    z, a, s, T, nu, cl = _atmosphere(37, 9, 5)
    if not cloud:
        cl = None
    r = raw_optics(_t(a), _t(s), _t(z), _t(T), _t(nu), None if cl is None else _t(cl), Ts, Tt)
    ref = oracle.raw_optics(a, s, z, T, nu, cloud_optical_depth=cl, surface_temperature=Ts,
                            top_temperature=Tt)
    for k, v in r.items():
        g = v.cpu().numpy()
        assert g.shape == ref[k].shape, k
        assert np.allclose(g, ref[k], rtol=4e-15, atol=0), (k, np.abs(g - ref[k]).max())


def test_raw_optics_clamped_layer_passes_nothing_back():
    W, L = 3, 4
    z, a, s, T, nu, cl = _order_one(W, L, 8, True)
    cl[1, :] = -50.0                                   # BOA-first layer 1: tau < 0, clamped
    args = [_t(a, True), _t(s, True), _t(z), _t(T), _t(nu), _t(cl, True)]
    r = raw_optics(*args)
    assert torch.all(r["delta_tau"][:, L - 2] == 0) and torch.all(r["single_scat_albedo"][:, L - 2] == 0)
    (r["delta_tau"].sum() + r["single_scat_albedo"].sum()).backward()
    assert torch.all(args[5].grad[1] == 0)
    assert torch.all(args[5].grad[0] != 0) and torch.all(args[5].grad[2] != 0)
    f = lambda a_, s_, c_: tuple(raw_optics(a_, s_, args[2], args[3], args[4], c_).values())
    assert torch.autograd.gradcheck(f, (args[0], args[1], args[5]), eps=1e-6, atol=1e-9,
                                    rtol=1e-7)


def _atmosphere(W, L, seed):
    """Level coefficients decaying with altitude, omega >= 0.05, thin clouds:
    stable doubling starts, so central differences of the oracle are smooth."""
    rng = np.random.default_rng(seed)
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(2e4, 8e5, L))])
    a = 10 ** rng.uniform(-8, -5.5, (L + 1, W)) * np.exp(-z / 8e5)[:, None]
    s = a * rng.uniform(0.1, 1.5, (L + 1, W))
    T = 300.0 - 70.0 * z / z[-1] + rng.normal(0, 2, L + 1)
    nu = np.sort(rng.uniform(300.0, 3000.0, W))
    cl = np.zeros((L, W))
    cl[L // 3] = rng.uniform(1e-4, 1e-2, W)
    return z, a, s, T, nu, cl


def _oracle_loss(a, s, z, T, nu, cl, Ts, Tt, mom, cfg, weights):
    r = oracle.raw_optics(a, s, z, T, nu, cloud_optical_depth=cl, surface_temperature=Ts,
                          top_temperature=Tt)
    o = oracle.solve_batch(r["delta_tau"], r["single_scat_albedo"], mom, cfg.num_quadrature,
                           planck_levels=r["planck_levels"], surface_emission=r["surface_emission"],
                           top_emission=r["top_emission"], surface_albedo=cfg.surface_albedo,
                           solar_flux=cfg.solar_flux, solar_mu=cfg.solar_mu,
                           use_delta_m=cfg.use_delta_m, extended=True)
    return sum(float(np.sum(weights[k] * o[k])) for k in OUTS)


@pytest.mark.parametrize("nq", [8, 16, 32])
def test_solve_from_coefficients_gradients_match_oracle_differences(nq):
    print(f"raw gradients N {nq}", flush=True)
    W, L = 4, 6
    z, a, s, T, nu, cl = _atmosphere(W, L, 30 + nq)
    rng = np.random.default_rng(nq)
    g = rng.uniform(0.3, 0.8, L)
    mom = g[:, None] ** np.arange(2 * nq + 4)
    cfg = BatchConfig(num_quadrature=nq, use_delta_m=True, surface_albedo=0.2, solar_flux=1.0,
                      solar_mu=0.6)
    Ts, Tt = 302.0, 150.0
    weights = {k: rng.uniform(0.5, 1.5, (W, L + 1)) for k in OUTS}
    full = nq <= 16
    ta, ts_, tT, tcl = _t(a, full), _t(s, full), _t(T, True), _t(cl, full)
    tTs, tTt = _t(Ts, True), _t(Tt, True)
    out = solve_from_coefficients(ta, ts_, _t(z), tT, _t(nu), _t(mom), cfg,
                                  cloud_optical_depth=tcl, surface_temperature=tTs,
                                  top_temperature=tTt)
    plain = solve_batch_from_coefficients(_t(a), _t(s), _t(z), _t(T), _t(nu), _t(mom), cfg,
                                          cloud_optical_depth=_t(cl), surface_temperature=Ts,
                                          top_temperature=Tt)
    for k in OUTS + ("flux_direct",):
        # same optics; at N <= 16 the solve records the reverse pass's
        # intermediates in kernels of their own, which round differently
        # (measured 1.6e-12): held to the FP64 parity bar
        x, y = out[k].detach().cpu().numpy(), plain[k].cpu().numpy()
        scale = np.abs(y).max(axis=1, keepdims=True)
        assert np.all(np.abs(x - y) <= 1e-10 * scale), k
    loss = sum((out[k] * _t(weights[k])).sum() for k in OUTS)
    loss.backward()

    base = dict(a=a, s=s, T=T, cl=cl, Ts=Ts, Tt=Tt)

    tau0 = oracle.raw_optics(a, s, z, T, nu, cloud_optical_depth=cl, surface_temperature=Ts,
                             top_temperature=Tt)["delta_tau"]

    def fd(name, idx):
        """Richardson-extrapolated central difference (error ~ h^4) of the
        extended-precision oracle, step 1e-3 of the entry (of the layer's
        optical depth for a cloud depth)."""
        x0 = base[name][idx] if idx is not None else base[name]
        h = 1e-3 * (tau0[idx[1], L - 1 - idx[0]] if name == "cl" else abs(x0))

        def at(step):
            v = {k: (np.array(x, dtype=np.float64) if isinstance(x, np.ndarray) else x)
                 for k, x in base.items()}
            if idx is not None:
                v[name][idx] += step
            else:
                v[name] += step
            return _oracle_loss(v["a"], v["s"], z, v["T"], nu, v["cl"], v["Ts"], v["Tt"], mom,
                                cfg, weights)
        d1 = (at(h) - at(-h)) / (2 * h)
        d2 = (at(2 * h) - at(-2 * h)) / (4 * h)
        return (4 * d1 - d2) / 3

    grads = {"T": tT.grad.cpu().numpy(), "Ts": float(tTs.grad), "Tt": float(tTt.grad)}
    picks = [("T", (0,)), ("T", (L // 2,)), ("T", (L,)), ("Ts", None), ("Tt", None)]
    if full:
        grads.update(a=ta.grad.cpu().numpy(), s=ts_.grad.cpu().numpy(), cl=tcl.grad.cpu().numpy())
        picks += [("a", (0, 1)), ("a", (L // 2, 2)), ("a", (L, 0)), ("s", (1, 3)),
                  ("s", (L - 1, 0)), ("cl", (L // 3, 1)), ("cl", (0, 2))]
    for name, idx in picks:
        num = fd(name, idx)
        ana = grads[name] if idx is None else grads[name][idx]
        scale = max(abs(num), abs(ana), 1e-300)
        assert abs(ana - num) <= 1e-7 * scale, (name, idx, ana, num)
