"""Newton-TGCR for the trapezoidal system on the GPU (drop-in for tdgl3d's
solvers/newton.py:76-120).

The whole Newton solve is one native call (``tdgl_newton_trap``): residual
g(x) = x - dt/2 f(x) - gamma, TGCR on (I - dt/2 J) dx = -g, the update and
the test ||g||_inf <= tol_f and ||dx||_inf <= tol_dx, with the reference's
return values (the last iterate, the converged flag, the iteration count).
``eval_f`` must be the device closure from ``integrators._make_eval_f``.
"""

from __future__ import annotations

import numpy as np

from .. import _native
from .tgcr import _device_closure


def newton_gcr_trap(eval_f, x0, gamma, dt: float, *, tol_f: float = 1e-3, tol_dx: float = 1e-3,
                    max_iter: int = 20, tol_gcr: float = 1e-4, eps_mf: float = 1e-4,
                    verbose: bool = False) -> tuple[np.ndarray, bool, int]:
    """(x, converged, iterations) of Newton from x0 on x - dt/2 f(x) = gamma."""
    ctx = _device_closure(eval_f, "newton_gcr_trap").bind()
    x = ctx.to_internal(np.asarray(x0))
    g = ctx.to_internal(np.asarray(gamma))
    opts = _native.solver_opts(tol_f, tol_dx, int(max_iter), tol_gcr, eps_mf)
    st = ctx.newton_trap(x, g, dt, opts)
    if verbose:
        print(f"  Newton-Trap: {st.newton_iters} iterations, ||g||inf = {st.g_inf:.3e}, "
              f"||dx||inf = {st.dx_inf:.3e}, TGCR {st.tgcr_iters_total} "
              f"({'converged' if st.converged else 'not converged'})")
    return ctx.to_reference(x).cpu().numpy(), bool(st.converged), int(st.newton_iters)
