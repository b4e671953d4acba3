"""Binding of one film (parameters, index maps, material) to a device context.

A :class:`FilmEngine` owns a libtdglhip context whose device buffers hold the
film's static data: hole-link bits, kappa and the superconducting mask.  All
compute goes through it; it never falls back to the host.
"""

from __future__ import annotations

import hashlib

import numpy as np

from . import _native


def _digest(*arrays) -> str:
    h = hashlib.blake2b(digest_size=16)
    for a in arrays:
        if a is None:
            h.update(b"none")
        else:
            a = np.ascontiguousarray(a)
            h.update(str(a.dtype).encode() + str(a.shape).encode())
            h.update(a.tobytes())
    return h.hexdigest()


class FilmEngine:
    def __init__(self, params, idx, material=None):
        self.params = params
        self.idx = idx
        self.material = material
        self.ctx = _native.Context(params.Nx, params.Ny, params.Nz, params.hx, params.hy,
                                   params.hz, params.kappa,
                                   (params.periodic_x, params.periodic_y, params.periodic_z))
        self._sig = None
        self._field = None
        self.refresh()

    def signature(self) -> str:
        p, m = self.params, self.material
        return _digest(np.array([p.Nx, p.Ny, p.Nz, p.hx, p.hy, p.hz, p.kappa, p.periodic_x,
                                 p.periodic_y, p.periodic_z], dtype=np.float64),
                       self.idx.hole_x_bc_mask, self.idx.hole_y_bc_mask, self.idx.hole_z_bc_mask,
                       None if m is None else m.kappa[self.idx.interior_to_full],
                       None if m is None else m.interior_sc_mask)

    def refresh(self) -> None:
        """Re-upload the static per-node data if it changed."""
        sig = self.signature()
        if sig == self._sig:
            return
        p, m = self.params, self.material
        bits = self.idx.hole_link_bits(p)
        if m is None:
            self.ctx.set_material(bits, None, None)
        else:
            self.ctx.set_material(bits, np.asarray(m.kappa)[self.idx.interior_to_full],
                                  np.asarray(m.interior_sc_mask, dtype=np.float64))
        self._sig = sig

    def set_boundary(self, u) -> None:
        """Load BoundaryVectors: three scalars when they came from
        build_boundary_field_vectors, the full vectors otherwise."""
        uni = getattr(u, "uniform", None)
        if uni is not None:
            if self._field != uni:
                self.ctx.set_field(*uni)
                self._field = uni
        else:
            self.ctx.set_field_vectors(u.Bx, u.By, u.Bz)
            self._field = None

    def set_field(self, bx: float, by: float, bz: float) -> None:
        uni = (float(bx), float(by), float(bz))
        if self._field != uni:
            self.ctx.set_field(*uni)
            self._field = uni


def engine_for(params, idx, material=None) -> FilmEngine:
    """Engine cached on the GridIndices object; rebuilt when the parameters or
    the material object change, refreshed when their contents change."""
    eng = getattr(idx, "_tdgl_engine", None)
    if eng is None or eng.params != params or eng.material is not material:
        if eng is not None:
            eng.ctx.close()
        eng = FilmEngine(params, idx, material)
        object.__setattr__(idx, "_tdgl_engine", eng)
    else:
        eng.refresh()
    return eng
