"""TDGL right-hand side on the GPU (drop-in for tdgl3d's physics/rhs.py:32-239).

``eval_f`` keeps the reference's signature and numpy-in/numpy-out contract;
the work is one launch of the fused stencil kernel (libtdglhip ``tdgl_rhs``),
which applies the boundary rules of ``_apply_boundary_conditions`` on the fly
instead of building ten CSR operators.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from ..core.material import MaterialMap
from ..core.parameters import SimulationParameters
from ..engine import engine_for
from ..mesh.indices import GridIndices


class BoundaryVectors:
    """(Bx, By, Bz) full-grid boundary vectors for one evaluation (rhs.py:32-40).
    ``uniform`` holds the scalar field when the vectors came from
    build_boundary_field_vectors."""

    __slots__ = ("Bx", "By", "Bz", "uniform")

    def __init__(self, Bx, By, Bz, uniform: Optional[tuple[float, float, float]] = None):
        self.Bx = Bx
        self.By = By
        self.Bz = Bz
        if uniform is None:
            uniform = getattr(Bx, "uniform", None)
        self.uniform = uniform

    @classmethod
    def from_field(cls, vecs) -> "BoundaryVectors":
        Bx, By, Bz = vecs
        return cls(Bx, By, Bz, uniform=getattr(vecs, "uniform", None))


def eval_f(X, params: SimulationParameters, idx: GridIndices, u: BoundaryVectors,
           material: Optional[MaterialMap] = None) -> np.ndarray:
    """dX/dt for the flat complex state X (rhs.py:147-239).  Returns a new
    complex128 array in the reference layout."""
    eng = engine_for(params, idx, material)
    eng.set_boundary(u)
    ctx = eng.ctx
    x = ctx.to_internal(np.asarray(X))
    f = ctx.empty_state()
    ctx.rhs(x, f)
    return ctx.to_reference(f).cpu().numpy()


class DeviceRHS:
    """The callable f(X) -> dX/dt that tdgl3d's Newton and TGCR consume
    (integrators.py:23-32), bound to one film and one set of boundary vectors.
    Called with a numpy state it returns eval_f's numpy result; the device
    drop-ins of ``newton_gcr_trap`` and ``tgcr_matrix_free_trap`` recognise it
    and run their whole solve in the library instead of calling it."""

    __slots__ = ("params", "idx", "u", "material")

    def __init__(self, params: SimulationParameters, idx: GridIndices, u: BoundaryVectors,
                 material: Optional[MaterialMap] = None):
        self.params = params
        self.idx = idx
        self.u = u
        self.material = material

    def __call__(self, X) -> np.ndarray:
        return eval_f(X, self.params, self.idx, self.u, self.material)

    def bind(self):
        """The film's device context with this closure's field loaded."""
        eng = engine_for(self.params, self.idx, self.material)
        eng.set_boundary(self.u)
        return eng.ctx


def device_observables(X, params, idx, u, material=None):
    """B = curl(phi) with the BCs applied and J_s without, at the interior
    nodes (the observables of core/solution.py:94-217), from the device."""
    eng = engine_for(params, idx, material)
    eng.set_boundary(u)
    ctx = eng.ctx
    x = ctx.to_internal(np.asarray(X))
    b, j = ctx.observables(x)
    n = params.n_interior
    b = b.cpu().numpy().reshape(3, n)
    j = j.cpu().numpy().reshape(3, n)
    return b, j
