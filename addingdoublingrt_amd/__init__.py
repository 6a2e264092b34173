"""MI355X-native batched adding-doubling radiative transfer (gfx950 HIP).

The compute path is ``libadrt_hip.so`` (C-ABI: include/adrt_hip.h); this
package is its host-side face:

* ``solve_batch`` / ``BatchConfig`` -- batched forward solve on device tensors
* ``solve_batch_from_coefficients`` -- the same from raw absorption/scattering
                                       coefficients, altitudes and temperatures
* ``adjoint_batch``                 -- hand-written reverse pass (gradients)
* ``solve``                         -- differentiable solve (torch.autograd)
* ``raw_optics`` / ``solve_from_coefficients``
                                    -- differentiable raw-coefficient optics and solve
* ``source_jacobians``              -- d outputs / d thermal sources (doubling once)
* ``planck_levels``                 -- device band Planck, differentiable in T
* ``workloads``                     -- synthetic inputs of the benchmark configs
"""
from ._lib import AdrtError, AdrtLibraryError, lib  # noqa: F401
from .solver import (BatchConfig, Workspace, adjoint_batch, gauss_legendre, planck,  # noqa: F401
                     planck_deriv, planck_levels, raw_optics, solve, solve_batch,
                     solve_batch_from_coefficients, solve_from_coefficients, source_jacobians)

__all__ = ["AdrtError", "AdrtLibraryError", "BatchConfig", "Workspace", "adjoint_batch",
           "gauss_legendre", "planck", "planck_deriv", "planck_levels", "raw_optics", "solve",
           "solve_batch", "solve_batch_from_coefficients", "solve_from_coefficients",
           "source_jacobians", "lib"]
