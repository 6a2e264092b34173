"""MI355X-native batched adding-doubling radiative transfer (gfx950 HIP).

The compute path is ``libadrt_hip.so`` (C-ABI: include/adrt_hip.h); this
package is its host-side face:

* ``solve_batch`` / ``BatchConfig`` -- batched columns on device tensors
* ``ADConfig`` / ``solve``          -- the per-column drop-in of the
  reference's ``adrt::ADConfig`` / ``adrt::solve``
* ``workloads``                     -- synthetic inputs of the benchmark configs
"""
from ._lib import AdrtError, AdrtLibraryError, lib  # noqa: F401
from .solver import BatchConfig, Workspace, gauss_legendre, planck, planck_deriv, solve_batch  # noqa: F401

__all__ = ["AdrtError", "AdrtLibraryError", "BatchConfig", "Workspace", "gauss_legendre",
           "planck", "planck_deriv", "solve_batch", "lib"]
