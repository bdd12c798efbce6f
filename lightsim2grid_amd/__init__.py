"""MI355X-native batched Newton-Raphson power flow behind lightsim2grid's solver surface.

The native module ``lightsim2grid_amd_cpp`` mirrors ``lightsim2grid.lightsim2grid_cpp``
for the Newton-Raphson path (SolverType, the solver classes, AnySolver, GridModel,
TimeSeriesCPP, ContingencyAnalysisCPP); its solvers run hand-written HIP kernels
for gfx950 through the C ABI in include/pfgpu.h. There is no CPU fallback:
importing the package without the built libraries raises.
"""
__version__ = "0.1.0"

__all__ = ["SolverType", "ErrorType", "solver", "compilation_options", "GridModel", "TimeSeriesCPP",
           "ContingencyAnalysisCPP", "gridgen", "abi"]

from .lightsim2grid_amd_cpp import SolverType, ErrorType  # noqa: E402
from .lightsim2grid_amd_cpp import GridModel, TimeSeriesCPP, ContingencyAnalysisCPP  # noqa: E402
from . import solver, compilation_options  # noqa: E402
