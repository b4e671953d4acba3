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

    def carve_spheres(self, centers, radii, params: SimulationParameters, idx) -> None:
        """Insulating spherical inclusions: sc_mask := 0 on every full-grid
        node within radius r (in cells) of a centre (in cells).  Extension of
        the reference's material API (which has polygons only); the hopping
        links stay, only the relaxation term of the insulator applies."""
        i = np.arange(params.Nx + 1, dtype=np.float64)
        j = np.arange(params.Ny + 1, dtype=np.float64)
        k = np.arange(params.Nz + 1 if params.is_3d else 1, dtype=np.float64)
        for (cx, cy, cz), r in zip(np.asarray(centers, float), np.asarray(radii, float)):
            i0, i1 = max(0, int(np.floor(cx - r))), min(params.Nx, int(np.ceil(cx + r)))
            j0, j1 = max(0, int(np.floor(cy - r))), min(params.Ny, int(np.ceil(cy + r)))
            k0 = max(0, int(np.floor(cz - r))) if params.is_3d else 0
            k1 = min(params.Nz, int(np.ceil(cz + r))) if params.is_3d else 0
            if i0 > i1 or j0 > j1 or k0 > k1:
                continue
            K, J, I = np.meshgrid(k[k0:k1 + 1], j[j0:j1 + 1], i[i0:i1 + 1], indexing="ij")
            inside = (I - cx) ** 2 + (J - cy) ** 2 + (K - cz) ** 2 <= r * r
            lin = I[inside].astype(np.int64) + params.mj * J[inside].astype(np.int64)
            if params.is_3d:
                lin = lin + params.mk * K[inside].astype(np.int64)
            self.sc_mask[lin] = 0.0
        self.interior_sc_mask = self.sc_mask[idx.interior_to_full]


def random_spheres(params: SimulationParameters, fraction: float, r_min: float, r_max: float,
                   seed: int, idx=None):
    """Centres uniform over the cell box and radii uniform in [r_min, r_max],
    drawn from default_rng(seed) until at least `fraction` of the interior
    nodes lie inside a sphere.  Returns (centres (m, 3), radii (m,))."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = params.Nx - 1, params.Ny - 1, (params.Nz - 1 if params.is_3d else 1)
    hit = np.zeros((nx, ny, nz), dtype=bool)
    ii = np.arange(1, params.Nx, dtype=np.float64)
    jj = np.arange(1, params.Ny, dtype=np.float64)
    kk = np.arange(1, params.Nz, dtype=np.float64) if params.is_3d else np.zeros(1)
    centers, radii = [], []
    target = fraction * hit.size
    while hit.sum() < target:
        c = rng.uniform([0.0, 0.0, 0.0], [params.Nx, params.Ny, params.Nz if params.is_3d else 0.0])
        r = rng.uniform(r_min, r_max)
        di = (ii - c[0]) ** 2
        dj = (jj - c[1]) ** 2
        dk = (kk - c[2]) ** 2
        inside = di[:, None, None] + dj[None, :, None] + dk[None, None, :] <= r * r
        hit |= inside
        centers.append(c)
        radii.append(r)
    return np.array(centers), np.array(radii)


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
