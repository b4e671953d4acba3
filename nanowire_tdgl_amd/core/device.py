"""Device = parameters + index maps + applied field + optional material
(mirror of tdgl3d's src/tdgl3d/core/device.py:18-169)."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from ..mesh.indices import GridIndices, construct_indices
from ..physics.applied_field import AppliedField
from .material import MaterialMap, Trilayer, build_material_map
from .parameters import SimulationParameters
from .state import StateVector


@dataclass
class Device:
    params: SimulationParameters
    applied_field: AppliedField = field(default_factory=AppliedField)
    trilayer: Optional[Trilayer] = None
    _idx: Optional[GridIndices] = field(default=None, repr=False, init=False)
    _material: Optional[MaterialMap] = field(default=None, repr=False, init=False)

    def __post_init__(self) -> None:
        if self.trilayer is not None and self.params.Nz != self.trilayer.Nz:
            object.__setattr__(self.params, "Nz", self.trilayer.Nz)
        self._idx = construct_indices(self.params)
        if self.trilayer is not None:
            self._material = build_material_map(self.params, self.trilayer, self._idx)

    @property
    def idx(self) -> GridIndices:
        if self._idx is None:
            self._idx = construct_indices(self.params)
        return self._idx

    @property
    def material(self) -> Optional[MaterialMap]:
        return self._material

    def initial_state(self) -> StateVector:
        """psi = 1 (times interior_sc_mask when a material exists), phi = 0."""
        sv = StateVector.uniform_superconducting(self.params)
        if self._material is not None:
            sv.psi[:] *= self._material.interior_sc_mask
        return sv

    def rebuild_indices(self) -> None:
        self._idx = construct_indices(self.params)
        if self.trilayer is not None:
            self._material = build_material_map(self.params, self.trilayer, self._idx)

    def add_hole(self, vertices, z_range: Optional[tuple[int, int]] = None) -> None:
        """Polygon hole with zero-current links; default z_range (0, Nz) includes
        both ghost planes (device.py:80-153).  Creates a uniform material map
        when none exists."""
        if z_range is None:
            z_range = (0, self.params.Nz)
        self.idx.define_hole_polygon(vertices, z_range, self.params)
        if self._material is None:
            p = self.params
            self._material = MaterialMap(kappa=np.full(p.dim_x, p.kappa, dtype=np.float64),
                                         sc_mask=np.ones(p.dim_x, dtype=np.float64),
                                         interior_sc_mask=np.ones(p.n_interior, dtype=np.float64))
        self._material.carve_hole_polygon(vertices, z_range, self.params, self.idx)

    def add_spheres(self, centers, radii) -> None:
        """Insulating spherical inclusions (centres and radii in cells) in the
        material's sc_mask; creates a uniform material map when none exists."""
        self._ensure_material()
        self._material.carve_spheres(centers, radii, self.params, self.idx)

    def _ensure_material(self) -> None:
        if self._material is None:
            p = self.params
            self._material = MaterialMap(kappa=np.full(p.dim_x, p.kappa, dtype=np.float64),
                                         sc_mask=np.ones(p.dim_x, dtype=np.float64),
                                         interior_sc_mask=np.ones(p.n_interior, dtype=np.float64))

    def __repr__(self) -> str:
        p = self.params
        tri = ""
        if self.trilayer is not None:
            r = self.trilayer.z_ranges()
            tri = (f", trilayer=({r['bottom'][1]}SC / {r['insulator'][1] - r['insulator'][0]}I / "
                   f"{r['top'][1] - r['top'][0]}SC)")
        f = self.applied_field
        return (f"Device(Nx={p.Nx}, Ny={p.Ny}, Nz={p.Nz}, κ={p.kappa}, "
                f"field=({f.Bx}, {f.By}, {f.Bz}){tri})")
