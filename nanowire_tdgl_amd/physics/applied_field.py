"""Applied magnetic field (mirror of tdgl3d's physics/applied_field.py:14-91)."""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from ..core.parameters import SimulationParameters


@dataclass
class AppliedField:
    """Uniform (Bx, By, Bz) on during (0, t_on_fraction * t_stop]; optional
    linear ramp over ramp_fraction * t_stop; or an arbitrary field_func(t, t_stop)."""

    Bx: float = 0.0
    By: float = 0.0
    Bz: float = 0.0
    t_on_fraction: float = 2.0 / 3.0
    ramp: bool = False
    ramp_fraction: float = 0.5
    field_func: Optional[Callable[[float, float], tuple[float, float, float]]] = None

    def evaluate(self, t: float, t_stop: float) -> tuple[float, float, float]:
        """(Bx, By, Bz) at time t (applied_field.py:42-56)."""
        if self.field_func is not None:
            return self.field_func(t, t_stop)
        if t <= 0 or t > t_stop * self.t_on_fraction:
            return 0.0, 0.0, 0.0
        scale = 1.0
        if self.ramp:
            t_ramp = t_stop * self.ramp_fraction
            scale = min(t / t_ramp, 1.0) if t_ramp > 0 else 1.0
        return self.Bx * scale, self.By * scale, self.Bz * scale


class FieldVectors(tuple):
    """The (Bx, By, Bz) full-grid vectors, remembering the uniform values they
    were built from so the device path can use three scalars instead."""

    uniform: tuple[float, float, float] | None = None


def build_boundary_field_vectors(applied_bx: float, applied_by: float, applied_bz: float,
                                 params: SimulationParameters, idx):
    """Full-grid vectors holding each component on the face-inner nodes of the
    two face pairs it enters the ghost rules through (applied_field.py:59-91)."""
    faces = {0: ("y", "z"), 1: ("z", "x"), 2: ("x", "y")}
    out = []
    for c, val in enumerate((applied_bx, applied_by, applied_bz)):
        v = np.zeros(params.dim_x, dtype=np.float64)
        if val != 0.0:
            a, b = faces[c]
            sel = np.concatenate([getattr(idx, f"{a}_face_lo_inner"), getattr(idx, f"{b}_face_lo_inner"),
                                  getattr(idx, f"{a}_face_hi_inner"), getattr(idx, f"{b}_face_hi_inner")])
            np.add.at(v, sel, val)
        out.append(v)
    res = FieldVectors(out)
    res.uniform = (float(applied_bx), float(applied_by), float(applied_bz))
    return res
