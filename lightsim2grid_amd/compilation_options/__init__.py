"""Compile-time feature flags, as lightsim2grid.compilation_options exposes them."""
__all__ = ["klu_solver_available", "nicslu_solver_available", "cktso_solver_available", "amdgpu_solver_available",
           "compiled_march_native", "compiled_o3_optim", "version"]

from ..lightsim2grid_amd_cpp import klu_solver_available  # noqa: E402
from ..lightsim2grid_amd_cpp import nicslu_solver_available  # noqa: E402
from ..lightsim2grid_amd_cpp import cktso_solver_available  # noqa: E402
from ..lightsim2grid_amd_cpp import amdgpu_solver_available  # noqa: E402
from ..lightsim2grid_amd_cpp import compiled_march_native  # noqa: E402
from ..lightsim2grid_amd_cpp import compiled_o3_optim  # noqa: E402
from ..lightsim2grid_amd_cpp import version  # noqa: E402
