"""Post-processing of saved states (mirror of tdgl3d's analysis package:
src/tdgl3d/analysis/__init__.py)."""

from __future__ import annotations

from .convergence import check_steady_state, compute_convergence_metrics
from .vortex_counting import (
    count_hole_flux_quanta,
    count_vortices_plaquette,
    count_vortices_polygon,
    find_vortex_cores,
)

__all__ = [
    "check_steady_state",
    "compute_convergence_metrics",
    "count_vortices_plaquette",
    "count_vortices_polygon",
    "count_hole_flux_quanta",
    "find_vortex_cores",
]
