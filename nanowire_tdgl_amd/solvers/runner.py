"""solve(): the user-facing entry point (drop-in for tdgl3d's
solvers/runner.py:24-163, same signature and return type)."""

from __future__ import annotations

from datetime import datetime
from pathlib import Path
from typing import Literal, Optional

import numpy as np

from ..core.device import Device
from ..core.solution import Solution
from ..core.state import StateVector
from ..io.logging import TimingContext, create_run_metadata
from ..physics.applied_field import AppliedField, build_boundary_field_vectors
from ..physics.rhs import BoundaryVectors
from .integrators import forward_euler, trapezoidal


def _make_eval_u(applied_field: AppliedField, params, idx, t_stop: float):
    """eval_u(t, X) -> BoundaryVectors of the applied field at t (runner.py:24-37)."""

    def eval_u(t, X):
        bx, by, bz = applied_field.evaluate(t, t_stop)
        return BoundaryVectors.from_field(build_boundary_field_vectors(bx, by, bz, params, idx))

    eval_u.ignores_state = True
    return eval_u


def solve(device: Device, t_start: float = 0.0, t_stop: float = 10.0, dt: float = 0.05,
          method: Literal["euler", "trapezoidal"] = "trapezoidal", x0=None, *,
          save_every: int = 1, progress: bool = True, verbose: bool = False,
          newton_tol_f: float = 1e-3, newton_tol_dx: float = 1e-3, newton_max_iter: int = 20,
          tol_gcr: float = 1e-4, eps_mf: float = 1e-4, adaptive: bool = True,
          log_metadata: bool = True, log_dir: str | Path = "logs",
          tgcr_max_iter: Optional[int] = None, precision: str = "f64",
          precond="chebyshev", predictor: str = "euler") -> Solution:
    """Run a TDGL simulation on the GPU (runner.py:40-163).  ``tgcr_max_iter``
    (None = the reference's n_state cap), ``precision`` ("f64", or "mixed":
    FP32 Krylov solve inside FP64 Newton), ``precond`` (the mixed solve's
    Chebyshev preconditioner: "chebyshev", "none" or a degree) and
    ``predictor`` ("euler": the reference's Newton start; "extrapolate": the
    extrapolated last step) are extensions."""
    params, idx = device.params, device.idx
    material = getattr(device, "material", None)
    if x0 is None:
        x0_arr = StateVector.uniform_superconducting(params).data
        if material is not None:
            sv = StateVector(x0_arr, params)
            sv.psi[:] *= material.interior_sc_mask
            x0_arr = sv.data
    elif isinstance(x0, StateVector):
        x0_arr = x0.data
    else:
        x0_arr = np.asarray(x0, dtype=np.complex128)

    eval_u = _make_eval_u(device.applied_field, params, idx, t_stop)
    log: list = []
    with TimingContext() as timer:
        if method == "euler":
            times, X_hist = forward_euler(x0_arr, params, idx, eval_u, t_start, t_stop, dt,
                                          save_every=save_every, progress=progress,
                                          material=material)
        elif method == "trapezoidal":
            times, X_hist = trapezoidal(x0_arr, params, idx, eval_u, t_start, t_stop, dt,
                                        newton_tol_f=newton_tol_f, newton_tol_dx=newton_tol_dx,
                                        newton_max_iter=newton_max_iter, tol_gcr=tol_gcr,
                                        eps_mf=eps_mf, save_every=save_every, adaptive=adaptive,
                                        progress=progress, verbose=verbose, material=material,
                                        tgcr_max_iter=tgcr_max_iter, step_log=log,
                                        precision=precision, precond=precond, predictor=predictor)
        else:
            raise ValueError(f"Unknown method {method!r}. Use 'euler' or 'trapezoidal'.")

    metadata_dict = None
    if log_metadata:
        extra = {}
        if log:
            extra = {"newton_iterations": int(sum(r["newton_iters"] for r in log)),
                     "dt_reductions": int(sum(1 for r in log if not r["converged"])),
                     "tgcr_iterations": int(sum(r["tgcr_iters_total"] for r in log))}
        meta = create_run_metadata(params=params, device=device, method=method, dt=dt,
                                   t_final=t_stop, wall_time=timer.elapsed,
                                   atol=newton_tol_f if method == "trapezoidal" else None,
                                   rtol=newton_tol_dx if method == "trapezoidal" else None,
                                   total_steps=len(times), **extra)
        metadata_dict = meta.to_dict()
        path = Path(log_dir)
        path.mkdir(parents=True, exist_ok=True)
        meta.save_json(path / f"run_{datetime.now().strftime('%Y%m%d_%H%M%S')}.json")
    return Solution(times=times, states=X_hist, params=params, idx=idx, device=device,
                    metadata=metadata_dict)
