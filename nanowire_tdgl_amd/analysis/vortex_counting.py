"""Vortex detection and flux counting on saved states (mirror of tdgl3d's
analysis/vortex_counting.py:14-431), vectorised over the plaquettes.

Results, orderings and edge behaviour are the reference's, including its
index conventions (interior states reshaped (nx, ny, nz) in the reference
interior order; the polygon fluxoid reads the full-grid vector reshaped
(Nx+1, Ny+1, Nz+1) in C order, as vortex_counting.py:216-235 does)."""

from __future__ import annotations

import numpy as np


def _wrap_phase(dphi):
    """Phase differences wrapped to [-pi, pi] (vortex_counting.py:434-436)."""
    return np.arctan2(np.sin(dphi), np.cos(dphi))


def _interior_shape(params):
    return params.Nx - 1, params.Ny - 1, max(params.Nz - 1, 1)


def count_vortices_plaquette(solution, device, slice_z: int = 0, step: int = -1,
                             winding_threshold: float = 0.8, mask_threshold: float = 0.02):
    """Phase winding around every elementary plaquette of an interior z-slice
    (vortex_counting.py:14-140).  Returns (count, centres (m, 2), winding numbers),
    in the reference's scan order (i outer, j inner)."""
    psi = solution.psi(step=step)
    nxi, nyi, nzi = _interior_shape(solution.params)
    psi_s = psi.reshape(nxi, nyi, nzi)[:, :, slice_z]
    p2 = np.abs(psi_s) ** 2
    ph = np.angle(psi_s)
    a, b, c, d = ph[:-1, :-1], ph[1:, :-1], ph[1:, 1:], ph[:-1, 1:]
    winding = ((_wrap_phase(b - a) + _wrap_phase(c - b)) + _wrap_phase(d - c)) + _wrap_phase(a - d)
    wn = winding / (2.0 * np.pi)
    masked = ((p2[:-1, :-1] < mask_threshold) | (p2[1:, :-1] < mask_threshold)
              | (p2[1:, 1:] < mask_threshold) | (p2[:-1, 1:] < mask_threshold))
    hit = (~masked) & (np.abs(wn) > winding_threshold)
    ii, jj = np.nonzero(hit)
    if ii.size == 0:
        return 0, np.empty((0, 2)), np.empty(0)
    pos = np.stack([ii + 0.5, jj + 0.5], axis=1).astype(np.float64)
    return int(ii.size), pos, wn[ii, jj]


def _expand(vals, params, idx):
    full = np.zeros(params.dim_x, dtype=np.asarray(vals).dtype)
    full[idx.interior_to_full] = vals
    return full


def count_vortices_polygon(solution, device, polygon_points, slice_z: int = 0, step: int = -1) -> float:
    """Phase winding along a polygon's vertices on a z-slice of the expanded
    state, in flux quanta (vortex_counting.py:143-273)."""
    p, idx = solution.params, solution.idx
    n = p.n_interior
    psi_full = _expand(solution.states[:n, step], p, idx)
    nx, ny = p.Nx + 1, p.Ny + 1
    if p.Nz == 1:
        sl = psi_full.reshape(nx, ny)
    else:
        sl = psi_full.reshape(nx, ny, p.Nz + 1)[:, :, slice_z + 1]
    poly = np.array(polygon_points)
    if not np.allclose(poly[0], poly[-1]):
        poly = np.vstack([poly, poly[0:1]])
    flux = 0.0
    for a, b in zip(poly[:-1], poly[1:]):
        i1, j1 = int(np.clip(a[0], 0, nx - 1)), int(np.clip(a[1], 0, ny - 1))
        i2, j2 = int(np.clip(b[0], 0, nx - 1)), int(np.clip(b[1], 0, ny - 1))
        flux += _wrap_phase(np.angle(sl[i2, j2]) - np.angle(sl[i1, j1]))
    return float(flux / (2.0 * np.pi))


def count_hole_flux_quanta(solution, device, hole_bounds, slice_z: int = 0, step: int = -1) -> float:
    """Magnetic flux sum(B_z) hx hy over an index box of an interior slice, in
    flux quanta 2 pi (vortex_counting.py:276-359); B_z from the device
    observable with the boundary conditions at times[step]."""
    p = solution.params
    _, _, bz = solution.bfield(step=step, full_interior=True)
    nxi, nyi, nzi = _interior_shape(p)
    bzs = np.asarray(bz).reshape(nxi, nyi, nzi)[:, :, slice_z]
    x0, x1, y0, y1 = hole_bounds
    i0, i1 = max(0, int(x0)), min(nxi, int(x1))
    j0, j1 = max(0, int(y0)), min(nyi, int(y1))
    flux = np.sum(bzs[i0:i1, j0:j1]) * p.hx * p.hy
    return float(flux / (2.0 * np.pi))


def find_vortex_cores(solution, device, slice_z: int = 0, step: int = -1, threshold: float = 0.1,
                      separation: int = 2):
    """Local minima of |psi|^2 below a threshold (vortex_counting.py:362-431)."""
    from scipy.ndimage import minimum_filter
    p2 = np.abs(solution.psi(step=step)) ** 2
    nxi, nyi, nzi = _interior_shape(solution.params)
    s = p2.reshape(nxi, nyi, nzi)[:, :, slice_z]
    is_min = (s == minimum_filter(s, size=separation)) & (s < threshold)
    return np.argwhere(is_min).astype(float)
