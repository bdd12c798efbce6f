"""Solver classes and enums, mirroring lightsim2grid.solver (lightsim2grid/solver/__init__.py).

The KLU / NICSLU / CKTSO / SparseLU classes of the reference are not part of this
build; their SolverType values exist and AnySolver/GridModel.change_solver raise
for them, as the reference does for a back-end that is not compiled.
"""
__all__ = ["SolverType", "ErrorType", "AnySolver", "AMDGPUSolver", "AMDGPUSolverSingleSlack", "AMDGPUMixedSolver",
           "AMDGPUMixedSolverSingleSlack", "AMDGPUDCSolver"]

from ..lightsim2grid_amd_cpp import SolverType  # noqa: E402
from ..lightsim2grid_amd_cpp import ErrorType  # noqa: E402
from ..lightsim2grid_amd_cpp import AnySolver  # noqa: E402
from ..lightsim2grid_amd_cpp import AMDGPUSolver  # SolverType.AMDGPU  # noqa: E402
from ..lightsim2grid_amd_cpp import AMDGPUSolverSingleSlack  # SolverType.AMDGPUSingleSlack  # noqa: E402
from ..lightsim2grid_amd_cpp import AMDGPUMixedSolver  # SolverType.AMDGPUMixed  # noqa: E402
from ..lightsim2grid_amd_cpp import AMDGPUMixedSolverSingleSlack  # SolverType.AMDGPUMixedSingleSlack  # noqa: E402
from ..lightsim2grid_amd_cpp import AMDGPUDCSolver  # SolverType.AMDGPUDC  # noqa: E402
