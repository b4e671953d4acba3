"""nanowire_tdgl_amd -- tdgl3d's trapezoidal TDGL time step on AMD MI355X.

Same call surface as tdgl3d (omedeiro/nanowire_tdgl, src/tdgl3d/__init__.py:9-46)
for the simulation path: parameters, device, state, material, applied field,
``solve``.  The right-hand side, the Newton/TGCR solve and the trapezoidal
update run in hand-written HIP kernels (libtdglhip.so); see include/tdgl_hip.h.
"""

from .core.device import Device
from .core.material import Layer, MaterialMap, Trilayer
from .core.parameters import SimulationParameters
from .core.solution import Solution
from .core.state import StateVector
from .physics.applied_field import AppliedField
from .solvers.runner import solve

# Analysis tools (tdgl3d re-exports the same six names, src/tdgl3d/__init__.py:17-25)
from .analysis import (
    check_steady_state,
    compute_convergence_metrics,
    count_hole_flux_quanta,
    count_vortices_plaquette,
    count_vortices_polygon,
    find_vortex_cores,
)

__version__ = "0.1.0"

__all__ = [
    "SimulationParameters",
    "Device",
    "StateVector",
    "Solution",
    "Layer",
    "Trilayer",
    "MaterialMap",
    "AppliedField",
    "solve",
    # Analysis
    "check_steady_state",
    "compute_convergence_metrics",
    "count_vortices_plaquette",
    "count_vortices_polygon",
    "count_hole_flux_quanta",
    "find_vortex_cores",
]
