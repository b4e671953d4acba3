"""Time integrators on the GPU (drop-in for tdgl3d's solvers/integrators.py:23-206).

The host loop keeps the reference's control flow exactly -- field frozen at
t_n for the implicit solve, dt/10 on Newton failure down to dt_min, dt
restored after an accepted step, history every `save_every` accepted steps
plus the final state -- while each step attempt is a single native call
(``tdgl_trap_step``: predictor, Newton, TGCR, all on the device).
"""

from __future__ import annotations

from typing import Callable, Optional

import numpy as np

from .. import _native
from ..core.material import MaterialMap
from ..core.parameters import SimulationParameters
from ..engine import engine_for
from ..mesh.indices import GridIndices
from ..physics.rhs import DeviceRHS


def _make_eval_f(params: SimulationParameters, idx: GridIndices, u,
                 material: Optional[MaterialMap] = None) -> DeviceRHS:
    """The closure f(X) -> dX/dt for the current boundary vectors
    (integrators.py:23-32); newton_gcr_trap and tgcr_matrix_free_trap run it
    on the device."""
    return DeviceRHS(params, idx, u, material)


def _bar(total, desc, enabled, unit="it"):
    if not enabled:
        return None
    try:
        from tqdm import tqdm
    except ImportError:
        return None
    return tqdm(total=total, desc=desc, unit=unit)


def _boundary(eval_u, t, eng, X):
    """eval_u(t, X) -> BoundaryVectors, handing the state over only to callables
    that read it (the ones built by solve() do not)."""
    if getattr(eval_u, "ignores_state", False):
        u = eval_u(t, None)
    else:
        u = eval_u(t, eng.ctx.to_reference(X).cpu().numpy())
    eng.set_boundary(u)


def forward_euler(x0, params: SimulationParameters, idx: GridIndices, eval_u: Callable, t_start: float,
                  t_stop: float, dt: float, *, save_every: int = 1, progress: bool = True,
                  material: Optional[MaterialMap] = None):
    """X += dt f(X) for ceil((t_stop - t_start)/dt) steps (integrators.py:35-89)."""
    eng = engine_for(params, idx, material)
    ctx = eng.ctx
    n_steps = int(np.ceil((t_stop - t_start) / dt))
    X0 = np.array(x0, dtype=np.complex128)
    X = ctx.to_internal(X0)
    times, hist = [t_start], [X0.copy()]
    t = t_start
    bar = _bar(n_steps, "Forward Euler", progress)
    for step in range(n_steps):
        dta = min(dt, t_stop - t)
        _boundary(eval_u, t, eng, X)
        ctx.euler_step(X, dta)
        t += dta
        if (step + 1) % save_every == 0 or step == n_steps - 1:
            times.append(t)
            hist.append(ctx.to_reference(X).cpu().numpy())
        if bar is not None:
            bar.update(1)
    if bar is not None:
        bar.close()
    return np.array(times), np.column_stack(hist)


def trapezoidal(x0, params: SimulationParameters, idx: GridIndices, eval_u: Callable, t_start: float,
                t_stop: float, dt: float, *, newton_tol_f: float = 1e-3, newton_tol_dx: float = 1e-3,
                newton_max_iter: int = 20, tol_gcr: float = 1e-4, eps_mf: float = 1e-4,
                save_every: int = 1, adaptive: bool = True, dt_min: float = 1e-8,
                progress: bool = True, verbose: bool = False,
                material: Optional[MaterialMap] = None, tgcr_max_iter: Optional[int] = None,
                step_log: Optional[list] = None, precision: str = "f64",
                mixed_tol_floor: float = 1e-6, precond="chebyshev", predictor: str = "euler"):
    """Implicit trapezoidal integration with Newton-TGCR (integrators.py:92-206).

    Extensions (defaults reproduce the reference): ``tgcr_max_iter`` caps the
    Krylov iterations per TGCR call (the reference always uses n_state);
    ``step_log`` receives one dict of native counters per step attempt;
    ``precision="mixed"`` keeps the Newton residual and the state in FP64 but
    runs TGCR with an analytic FP32 Jacobian-vector product and FP32 Krylov
    vectors, with its tolerance floored at ``mixed_tol_floor`` and (``precond``)
    a Chebyshev polynomial right preconditioner: "chebyshev" (automatic
    degree), "none", or a degree d >= 1.  ``predictor="extrapolate"`` starts
    each Newton iteration after the first accepted step from the linear
    extrapolation X + dt/dt_prev (X - X_prev) of the last accepted step instead
    of the reference's X + dt f(X) (integrators.py:156); the step equations
    and the convergence test are the same.
    """
    if predictor not in ("euler", "extrapolate"):
        raise ValueError("predictor must be 'euler' or 'extrapolate'")
    eng = engine_for(params, idx, material)
    ctx = eng.ctx
    X0 = np.array(x0, dtype=np.complex128)
    X = ctx.to_internal(X0)
    opts = _native.solver_opts(newton_tol_f, newton_tol_dx, newton_max_iter, tol_gcr, eps_mf,
                               tgcr_max_iter, precision=precision,
                               mixed_tol_floor=mixed_tol_floor, precond=precond)
    t = t_start
    cur = dt
    times, hist = [t], [X0.copy()]
    step = 0
    bar = _bar(t_stop - t_start, "Trapezoidal", progress, unit="t")
    x_prev, dt_prev = None, 0.0
    while t < t_stop - 1e-14:
        dta = min(cur, t_stop - t)
        _boundary(eval_u, t, eng, X)
        x_in = X.clone() if predictor == "extrapolate" else None
        st = ctx.trap_step(X, dta, opts, x_prev=x_prev, dt_prev=dt_prev)
        if step_log is not None:
            rec = st.as_dict()
            rec.update(t=t, dt=dta)
            step_log.append(rec)
        if verbose:
            print(f"  t={t:.6e} dt={dta:.3e} newton={st.newton_iters} tgcr={st.tgcr_iters_total} "
                  f"||g||inf={st.g_inf:.3e} ||dx||inf={st.dx_inf:.3e} "
                  f"{'ok' if st.converged else 'FAILED'}")
        if st.converged:
            t += dta
            step += 1
            if x_in is not None:
                x_prev, dt_prev = x_in, dta
            if bar is not None:
                bar.update(dta)
            if step % save_every == 0:
                times.append(t)
                hist.append(ctx.to_reference(X).cpu().numpy())
            cur = dt
        else:
            if not adaptive:
                raise RuntimeError(f"Newton did not converge at t={t:.6e} with dt={dta:.3e}.")
            cur /= 10.0
            if cur < dt_min:
                raise RuntimeError(f"dt reduced below dt_min={dt_min:.1e} at t={t:.6e}; aborting.")
            if verbose:
                print(f"  dt reduced to {cur:.3e}")
    if bar is not None:
        bar.close()
    if times[-1] < t:
        times.append(t)
        hist.append(ctx.to_reference(X).cpu().numpy())
    return np.array(times), np.column_stack(hist)
