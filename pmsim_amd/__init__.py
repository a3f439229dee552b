"""MI355X-native implicit nonlinear solid element loop (pmsim's solid / von Mises path).

Host-side mirror of pmsim's builder API; the compute path is the C-ABI library
``libpmsim_amd.so`` (HIP kernels for gfx950), loaded on first use by ``pmsim_amd.solver``.
"""
from .mesh import At, Features, Mesh, any_x, points_of  # noqa: F401
from .problem import (Config, Dof, Elem, Essential, FemBase, Natural, Nbc, ParamSolid, Pbc,  # noqa: F401
                      Settings, StressStrain, describe)

__all__ = ["At", "Features", "Mesh", "any_x", "points_of", "Config", "Dof", "Elem", "Essential", "FemBase",
           "Natural", "Nbc", "ParamSolid", "Pbc", "Settings", "StressStrain", "describe"]
