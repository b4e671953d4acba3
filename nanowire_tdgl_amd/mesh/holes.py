"""Hole geometry on the full grid (mirror of tdgl3d's mesh/holes.py:43-311).

The node and link masks must equal the reference's bit for bit, including
array order, so the even-odd ray cast uses the same half-open comparisons and
the same floating-point expression for the edge crossing, vectorised over all
nodes of a z-column at once.
"""

from __future__ import annotations

from typing import Literal

import numpy as np


def point_in_polygon(point: tuple[float, float], vertices: list[tuple[float, float]]) -> bool:
    """Even-odd test of one point (holes.py:43-100)."""
    inside = _ray_cast(np.array([float(point[0])]), np.array([float(point[1])]), vertices)
    return bool(inside[0])


def _ray_cast(x: np.ndarray, y: np.ndarray, vertices) -> np.ndarray:
    """Vectorised even-odd ray cast toward +x.  An edge (a, b) toggles a point
    when min(ay, by) < y <= max(ay, by), x <= max(ax, bx) and the point is left
    of the crossing (or the edge is vertical).  Horizontal edges never pass the
    first test, so the crossing is always well defined where it is used."""
    inside = np.zeros(x.shape, dtype=bool)
    nv = len(vertices)
    for s in range(nv):
        ax, ay = (float(v) for v in vertices[s])
        bx, by = (float(v) for v in vertices[(s + 1) % nv])
        hit = (y > min(ay, by)) & (y <= max(ay, by)) & (x <= max(ax, bx))
        if not hit.any():
            continue
        if ax == bx:
            inside ^= hit
            continue
        if ay != by:
            with np.errstate(invalid="ignore", divide="ignore"):
                xc = (y - ay) * (bx - ax) / (by - ay) + ax
            inside ^= hit & (x <= xc)
    return inside


def identify_hole_nodes(vertices, z_range, grid_spacing_x, grid_spacing_y, Nx, Ny, Nz) -> np.ndarray:
    """Boolean (Nx+1, Ny+1, Nz+1) mask of nodes inside the polygon, extruded
    over k in [z_min, min(z_max + 1, Nz + 1)) (holes.py:103-157)."""
    ii, jj = np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), indexing="ij")
    xs = ii.astype(np.float64) * grid_spacing_x
    ys = jj.astype(np.float64) * grid_spacing_y
    # i * h with a Python int and float rounds like the int->double cast then multiply
    inside2d = _ray_cast(xs.ravel(), ys.ravel(), vertices).reshape(Nx + 1, Ny + 1)
    mask = np.zeros((Nx + 1, Ny + 1, Nz + 1), dtype=bool)
    k0, k1 = int(z_range[0]), int(z_range[1])
    ks = [k for k in range(k0, min(k1 + 1, Nz + 1))]
    if ks:
        mask[:, :, ks] = inside2d[:, :, None]
    return mask


def identify_boundary_links(hole_mask: np.ndarray, direction: Literal["x", "y", "z"],
                            is_3d: bool = True) -> np.ndarray:
    """Full-grid ids of links with exactly one end in the hole, in the
    reference's loop order: x -> (k, j, i), y -> (k, i, j), z -> (i, j, k)
    (holes.py:160-257)."""
    Nx, Ny, Nz = (s - 1 for s in hole_mask.shape)
    mj, mk = Nx + 1, (Nx + 1) * (Ny + 1)
    if direction == "x":
        d = hole_mask[:-1, :, :] != hole_mask[1:, :, :]          # (i, j, k)
        k, j, i = np.nonzero(d.transpose(2, 1, 0))
    elif direction == "y":
        d = hole_mask[:, :-1, :] != hole_mask[:, 1:, :]
        k, i, j = np.nonzero(d.transpose(2, 0, 1))
    elif direction == "z":
        d = hole_mask[:, :, :-1] != hole_mask[:, :, 1:]
        i, j, k = np.nonzero(d)
    else:
        raise ValueError(f"Invalid direction '{direction}'. Use 'x', 'y', or 'z'.")
    m = i.astype(np.int64) + mj * j.astype(np.int64)
    if is_3d:
        m = m + mk * k.astype(np.int64)
    return m.astype(np.int64)


def identify_circular_hole_nodes(center, radius, z_range, grid_spacing_x, grid_spacing_y,
                                 Nx, Ny, Nz) -> np.ndarray:
    """Nodes strictly inside a circle (holes.py:260-311)."""
    ii, jj = np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), indexing="ij")
    x = ii * float(grid_spacing_x)
    y = jj * float(grid_spacing_y)
    dist = np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2)
    mask = np.zeros((Nx + 1, Ny + 1, Nz + 1), dtype=bool)
    ks = list(range(int(z_range[0]), min(int(z_range[1]) + 1, Nz + 1)))
    if ks:
        mask[:, :, ks] = (dist < radius)[:, :, None]
    return mask
