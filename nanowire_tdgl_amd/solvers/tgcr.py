"""Matrix-free TGCR for the trapezoidal system on the GPU (drop-in for
tdgl3d's solvers/tgcr.py:96-156).

The whole solve is one native call (``tdgl_tgcr_trap``): the finite-difference
product, modified Gram-Schmidt, the breakdown test and the update run on the
device with the reference's control flow, so the iteration count and the
empty-array failure are the reference's.  ``eval_f`` must be the device
closure that ``integrators._make_eval_f`` returns; any other callable is
rejected (there is no host fallback).
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from .. import _native
from ..physics.rhs import DeviceRHS


def _device_closure(eval_f, who: str) -> DeviceRHS:
    if not isinstance(eval_f, DeviceRHS):
        raise TypeError(f"{who} runs on the device and needs the closure from "
                        "nanowire_tdgl_amd.solvers.integrators._make_eval_f, not "
                        f"{type(eval_f).__name__}")
    return eval_f


def tgcr_matrix_free_trap(eval_f, x_lin, b, dt: float, tol: float = 1e-4,
                          max_iter: Optional[int] = None, eps_mf: float = 1e-4) -> np.ndarray:
    """dx with (I - dt/2 J(x_lin)) dx = b; an empty complex128 array when TGCR
    does not reach ``tol`` (tgcr.py:154-155).  Arrays in the reference layout."""
    ctx = _device_closure(eval_f, "tgcr_matrix_free_trap").bind()
    xl = ctx.to_internal(np.asarray(x_lin))
    bb = ctx.to_internal(np.asarray(b))
    if max_iter is not None and max_iter <= 0:
        # the loop never runs (tgcr.py:124): x = 0 stands only for b = 0
        if ctx.dot(bb, bb) == 0.0:
            return np.zeros(len(b), dtype=np.complex128)
        return np.array([], dtype=np.complex128)
    opts = _native.solver_opts(tol_gcr=tol, eps_mf=eps_mf,
                               tgcr_max_iter=None if max_iter is None else int(max_iter))
    dx = ctx.empty_state()
    ok, _ = ctx.tgcr_trap(xl, bb, dt, opts, dx)
    if not ok:
        return np.array([], dtype=np.complex128)
    return ctx.to_reference(dx).cpu().numpy()
