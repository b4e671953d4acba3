"""Grid and physics parameters (mirror of tdgl3d's SimulationParameters,
src/tdgl3d/core/parameters.py:9-94: same fields, defaults, validation and
derived sizes)."""

from __future__ import annotations

import copy
from dataclasses import dataclass


@dataclass
class SimulationParameters:
    """Nx, Ny, Nz count cells; the full grid has (Nx+1)(Ny+1)(Nz+1) nodes and
    the PDE lives on the (Nx-1)(Ny-1)(Nz-1) interior nodes.  Nz == 1 is the
    quasi-2-D film."""

    Nx: int = 10
    Ny: int = 10
    Nz: int = 1
    hx: float = 1.0
    hy: float = 1.0
    hz: float = 1.0
    kappa: float = 5.0
    periodic_x: bool = False
    periodic_y: bool = False
    periodic_z: bool = False

    def __post_init__(self) -> None:
        if self.Nx < 2 or self.Ny < 2:
            raise ValueError("Nx and Ny must each be >= 2.")
        if self.Nz < 1:
            raise ValueError("Nz must be >= 1.")

    @property
    def is_3d(self) -> bool:
        return self.Nz > 1

    @property
    def n_interior(self) -> int:
        return (self.Nx - 1) * (self.Ny - 1) * max(self.Nz - 1, 1)

    @property
    def n_full(self) -> int:
        planes = self.Nz + 1 if self.is_3d else 1
        return (self.Nx + 1) * (self.Ny + 1) * planes

    @property
    def n_state(self) -> int:
        return (4 if self.is_3d else 3) * self.n_interior

    @property
    def dim_x(self) -> int:
        return self.n_full

    @property
    def mj(self) -> int:
        return self.Nx + 1

    @property
    def mk(self) -> int:
        return (self.Nx + 1) * (self.Ny + 1)

    def copy(self) -> "SimulationParameters":
        return copy.deepcopy(self)
