"""MI355X-native DIABLO duct time step (RK3 + Crank-Nicolson ADI + per-kx
multigrid pressure projection) behind the reference's subroutine surface.

The compute path is the HIP library ``libekman_bl_amd.so`` built from
``ekman_bl_amd/csrc``; `DuctSolver` is its Python binding.  Importing the
package does not load the library; constructing a solver does, and fails
loudly when it is missing.
"""
from .params import EkmParams, config, base_params  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    if name in ("DuctSolver", "EkmDiag"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
