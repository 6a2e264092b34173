"""The dynamic schedules of the solve, through the C-ABI.

Two parts of the solve hand out work at run time: the prep kernel appends the
columns that take the reference's composite-form adding (two unstable
scattering layers in a row) to a list by atomic increments, in no set order,
and the composite kernels take their columns from it; the persistent grids of
the reverse doubling take layers from a task counter, so which wave (and which
checkpoint slab) runs a layer changes from run to run.  Neither may change a
result: each column's adding and each layer's reverse are computed the same way
whichever wave runs them.  Checked bit for bit, on batches with flagged columns
spread over the whole batch: two runs of the same workspace, and a workspace
that splits the batch into several chunks (the list and the counters rebuilt
per chunk) against one that holds it whole, at orders 8 (pair mode on the
reverse-pass workspace), 16 (FP64 and the mixed_reverse form) and 32.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from addingdoublingrt_amd import BatchConfig, adjoint_batch, solve_batch  # noqa: E402
from addingdoublingrt_amd.solver import Workspace  # noqa: E402
from test_gpu_adjoint import random_cot, t, unstable_layers  # noqa: E402

pytestmark = pytest.mark.gpu


def flagged_batch(n, W, L, seed):
    rng = np.random.default_rng(seed)
    tau = 10.0 ** rng.uniform(-2.0, 1.0, (W, L))
    ssa = rng.uniform(0.15, 0.97, (W, L))
    g = rng.uniform(0.3, 0.85, (W, L))
    mom = g[..., None] ** np.arange(2 * n)
    B = rng.uniform(0.5, 2.0, (W, L + 1))
    for w in range(0, W, 3):  # every third column: an unstable pair
        ssa[w, 2], ssa[w, 3] = 0.005, 0.004
        tau[w, 2], tau[w, 3] = 0.3, 0.2
    return tau, ssa, mom, B


def run(args, kw, cot, precision, max_bytes, generation=True):
    out = solve_batch(*args, **kw, adjoint_workspace=True, max_workspace_bytes=max_bytes)
    gen = out.pop("_generation")
    g = adjoint_batch(*args, out, cot, **kw, precision=precision,
                      forward_generation=gen if generation else 0, max_workspace_bytes=max_bytes)
    torch.cuda.synchronize()
    return out, g


def assert_identical(a, b, label):
    for k in a:
        if a[k] is None:
            continue
        assert torch.equal(a[k], b[k]), f"{label}: {k}"


@pytest.mark.parametrize("n,precision", [(8, "f64"), (16, "f64"), (16, "mixed_reverse"),
                                         (32, "f64")])
def test_dynamic_schedules_are_deterministic(n, precision):
    W, L = 96, 10
    tau, ssa, mom, B = flagged_batch(n, W, L, seed=40 + n)
    cfg = BatchConfig(num_quadrature=n, surface_albedo=0.2, solar_flux=1.0, solar_mu=0.6)
    loose = unstable_layers(tau, ssa, mom, cfg)
    assert loose[::3, 2].all() and loose[::3, 3].all()
    args = (t(tau), t(ssa), t(mom), cfg)
    kw = dict(planck_levels=t(B), surface_emission=t(B[:, -1] * 1.1),
              top_emission=t(B[:, 0] * 0.9))
    cot = {k: t(v) for k, v in random_cot(W, L, 9, ("flux_up", "flux_down", "mean_intensity",
                                                   "flux_direct")).items()}
    out1, g1 = run(args, kw, cot, precision, 0)
    out2, g2 = run(args, kw, cot, precision, 0)
    assert_identical(out1, out2, "forward, repeated")
    assert_identical(g1, g2, "gradients, repeated")
    # a budget that holds about a quarter of the batch per chunk
    small = Workspace(W // 4, L, n, 2 * n, adjoint=True).bytes + 4096
    ws = Workspace(W, L, n, 2 * n, max_bytes=small, adjoint=True)
    assert 1 < ws.chunk < W, ws.chunk
    del ws
    out3, g3 = run(args, kw, cot, precision, small, generation=False)
    assert_identical(out1, out3, "forward, chunked")
    assert_identical(g1, g3, "gradients, chunked")
