"""Per-node material maps and layer stacks (mirror of tdgl3d's
src/tdgl3d/core/material.py:26-254)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .parameters import SimulationParameters


@dataclass
class MaterialMap:
    """kappa and sc_mask per full-grid node; interior_sc_mask per interior node
    (interior order).  sc_mask = 1 superconductor, 0 insulator/hole."""

    kappa: np.ndarray
    sc_mask: np.ndarray
    interior_sc_mask: np.ndarray

    def carve_hole_polygon(self, vertices, z_range, params: SimulationParameters, idx) -> None:
        """sc_mask := 0 on the polygon's nodes; interior copy refreshed (material.py:43-112)."""
        from ..mesh.holes import identify_hole_nodes

        mask = identify_hole_nodes(vertices, z_range, params.hx, params.hy,
                                   params.Nx, params.Ny, params.Nz)
        ii, jj, kk = np.nonzero(mask)
        lin = ii + params.mj * jj
        if params.is_3d:
            lin = lin + params.mk * kk
        self.sc_mask[lin] = 0.0
        self.interior_sc_mask = self.sc_mask[idx.interior_to_full]


@dataclass
class Layer:
    thickness_z: int
    kappa: float
    is_superconductor: bool = True


@dataclass
class Trilayer:
    """Superconductor / insulator / superconductor stack along z from k = 0."""

    bottom: Layer
    insulator: Layer
    top: Layer

    def __post_init__(self) -> None:
        if self.insulator.is_superconductor:
            raise ValueError("The insulator layer must have is_superconductor=False.")

    @property
    def Nz(self) -> int:
        return self.bottom.thickness_z + self.insulator.thickness_z + self.top.thickness_z

    def z_ranges(self) -> dict[str, tuple[int, int]]:
        b, i = self.bottom.thickness_z, self.insulator.thickness_z
        return {"bottom": (0, b), "insulator": (b, b + i),
                "top": (b + i, b + i + self.top.thickness_z)}


def build_material_map(params: SimulationParameters, trilayer: Trilayer, idx) -> MaterialMap:
    """Per-z-plane kappa / sc from the stack; plane k belongs to the layer whose
    cell range [start, end) holds k, and k == Nz to the top (material.py:182-254)."""
    if params.Nz != trilayer.Nz:
        raise ValueError(f"params.Nz={params.Nz} does not match trilayer.Nz={trilayer.Nz}")
    layers = {"bottom": trilayer.bottom, "insulator": trilayer.insulator, "top": trilayer.top}
    ranges = trilayer.z_ranges()
    plane = params.mk
    kappa = np.empty(params.dim_x, dtype=np.float64)
    sc = np.empty(params.dim_x, dtype=np.float64)
    for k in range(params.Nz + 1):
        name = next((nm for nm, (lo, hi) in ranges.items() if lo <= k < hi), "top")
        lay = layers[name]
        kappa[k * plane:(k + 1) * plane] = lay.kappa
        sc[k * plane:(k + 1) * plane] = 1.0 if lay.is_superconductor else 0.0
    return MaterialMap(kappa=kappa, sc_mask=sc, interior_sc_mask=sc[idx.interior_to_full])
