"""Gradients through the trapezoidal integrator: a hand-written discrete adjoint.

The forward model is the plain FP64 trapezoidal method (integrators.py:92-206):
each step solves

    G(x_{n+1}) = x_{n+1} - x_n - dt/2 (F(x_{n+1}; u_n) + F(x_n; u_n)) = 0

with the field u_n of the step's start time.  For a loss L(x_N, u_N) the
discrete adjoint runs backwards:

    (I - dt/2 J(x_{n+1}))^T mu = lambda_{n+1}
    lambda_n = mu + dt/2 J(x_n)^T mu
    dL/dtheta += dt/2 mu^T (dF/dtheta(x_{n+1}) + dF/dtheta(x_n))

with J^T w and the parameter derivatives from the library's adjoint kernels
(``tdgl_vjp_f``, ``tdgl_vjp_bz``); the linear adjoint solves are the library's
TGCR on the transposed step operator (``tdgl_adjoint_solve``, FP64, with the
mixed method's Chebyshev right preconditioner by default).  States are checkpointed every ``checkpoint_every`` steps in the
forward pass and each segment is recomputed (bit for bit: the FP64 path is
deterministic) before its adjoint sweep.

Parameters: kappa (every node's kappa tied), the applied-field amplitude Bz
(the field of step n is Bz * s(t_n), s the AppliedField time profile), and a
per-node relaxation/pinning field w with sc_mask = 1 - w.  The result is exact
for the discrete solution; Newton stops at its tolerance, so compare with
finite differences of runs solved tightly.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native
from .core.device import Device
from .core.material import MaterialMap
from .core.parameters import SimulationParameters
from .engine import FilmEngine
from .physics.applied_field import AppliedField


def _internal_to_reference(params: SimulationParameters) -> np.ndarray:
    """perm[q] = reference interior index of internal node q (i fastest vs k fastest)."""
    nx, ny = params.Nx - 1, params.Ny - 1
    nz = params.Nz - 1 if params.Nz > 1 else 1
    q = np.arange(nx * ny * nz)
    i, t = q % nx, q // nx
    j, k = t % ny, t // ny
    return (i * ny + j) * nz + k


@dataclass
class AdjointProblem:
    """A film, its initial state and the time stepping of the gradient study."""

    params: SimulationParameters
    x0: np.ndarray                          # reference layout, complex128
    dt: float = 0.2
    steps: int = 400
    checkpoint_every: int = 20
    t_on_fraction: float = 1.0
    holes: list = field(default_factory=list)
    # Tight Newton; TGCR at the reference's default tolerance, because its
    # absolute breakdown test (||Ap|| < 1e-14, tgcr.py:141-142) ends a solve
    # whose target tol * ||r0|| falls below 1e-14 near convergence.
    newton_tol_f: float = 1e-10
    newton_tol_dx: float = 1e-8
    tol_gcr: float = 1e-4
    adjoint_tol: float = 1e-12
    adjoint_precond: object = "chebyshev"  # the adjoint TGCR's preconditioner (as the mixed method's)
    # The forward steps' FP64 TGCR: "chebyshev" (the same polynomial as a right
    # preconditioner, the finite-difference JVP in its recurrence; an FP64 step
    # that converges to the same solution with far fewer Krylov iterations) or
    # "none" (the reference's TGCR).
    forward_precond: object = "chebyshev"
    # Newton's start in the forward steps: "extrapolate" (X + (X - X_prev),
    # fewer Newton iterations) or "euler" (X + dt f, the reference's); with the
    # tight Newton tolerances above both give the same trajectory to ~1e-10.
    # Checkpoints keep X_prev, so the recomputed segments repeat the forward
    # steps bit for bit either way.
    predictor: str = "extrapolate"

    def __post_init__(self):
        self.device = Device(self.params)
        for verts, zr in self.holes:
            self.device.add_hole(verts, zr)
        self.device._ensure_material()
        self.material: MaterialMap = self.device.material
        self._base_sc = self.material.sc_mask.copy()
        self.engine = FilmEngine(self.params, self.device.idx, self.material)
        self.ctx = self.engine.ctx
        self.t_stop = self.steps * self.dt
        self._perm = _internal_to_reference(self.params)
        self.stats: dict = {}
        # forward runs so far: backward() needs the checkpoints of the run it differentiates
        self.generation = 0

    # -- parameters -----------------------------------------------------------
    def set_parameters(self, kappa: float, w_full: np.ndarray) -> None:
        """kappa for every node; sc_mask = base mask (holes) * (1 - w)."""
        m = self.material
        m.kappa[:] = float(kappa)
        m.sc_mask[:] = self._base_sc * (1.0 - np.asarray(w_full, dtype=np.float64))
        m.interior_sc_mask = m.sc_mask[self.device.idx.interior_to_full]
        self.params.kappa = float(kappa)
        self.engine.refresh()

    def profile(self, t: float) -> float:
        """s(t): the field of the step that starts at t is Bz * s(t)."""
        return AppliedField(Bz=1.0, t_on_fraction=self.t_on_fraction).evaluate(t, self.t_stop)[2]

    def _set_field(self, bz: float, t: float) -> None:
        self.ctx.set_field(0.0, 0.0, bz * self.profile(t))

    # -- forward --------------------------------------------------------------
    def _step(self, X, bz: float, t: float, x_prev=None) -> None:
        self._set_field(bz, t)
        st = self.ctx.trap_step(X, self.dt, self._opts, x_prev=x_prev, dt_prev=self.dt)
        if not st.converged:
            raise RuntimeError(f"trapezoidal step at t={t:.4g} did not converge at dt={self.dt} "
                               "(the adjoint assumes a fixed step; loosen the tolerances)")
        self.stats["newton"] = self.stats.get("newton", 0) + st.newton_iters
        self.stats["tgcr"] = self.stats.get("tgcr", 0) + st.tgcr_iters_total

    def forward(self, kappa: float, bz: float, w_full: np.ndarray):
        """Run the steps; returns (final internal state, B_z observable (reference order))."""
        import torch
        # a new run: outputs of earlier runs can no longer be differentiated
        self.generation += 1
        self.set_parameters(kappa, w_full)
        self._opts = _native.solver_opts(self.newton_tol_f, self.newton_tol_dx, 20, self.tol_gcr,
                                         f64_precond=self.forward_precond)
        if self.predictor not in ("euler", "extrapolate"):
            raise ValueError("predictor must be 'euler' or 'extrapolate'")
        ex = self.predictor == "extrapolate"
        X = self.ctx.to_internal(self.x0)
        # checkpoint n: (X_n, X_{n-1}) -- the second only for the extrapolating predictor
        self.checkpoints = {0: (X.clone(), None)}
        prev = None
        for n in range(self.steps):
            x_in = X.clone() if ex else None
            self._step(X, bz, n * self.dt, prev)
            prev = x_in
            if (n + 1) % self.checkpoint_every == 0:
                self.checkpoints[n + 1] = (X.clone(), prev)
        self.X_final = X
        self.ctx.set_field(0.0, 0.0, bz * self.profile(self.t_stop))
        b, _ = self.ctx.observables(X, want_b=True, want_j=False)
        n = self.params.n_interior
        self._bz_param = bz
        return X, b[2 * n:3 * n].clone()

    # -- backward -------------------------------------------------------------
    def backward(self, g_bz):
        """Gradients of <g_bz, B_z(x_N)> w.r.t. (kappa, Bz, w_full)."""
        import torch
        ctx, dt = self.ctx, self.dt
        hdt = dt / 2.0
        bz = self._bz_param
        n = self.params.n_interior
        lam = ctx.empty_state()
        ctx.set_field(0.0, 0.0, bz * self.profile(self.t_stop))
        df = ctx.vjp_bz(g_bz.contiguous(), lam, field=True)
        d_bz = df[2] * self.profile(self.t_stop)
        d_kappa = 0.0
        dsc = torch.zeros(n, dtype=torch.float64, device=ctx.device)
        tmp = torch.empty(n, dtype=torch.float64, device=ctx.device)
        g = ctx.empty_state()
        iters = 0
        marks = sorted(self.checkpoints)
        for s in range(len(marks) - 1, -1, -1):
            lo = marks[s]
            hi = marks[s + 1] if s + 1 < len(marks) else self.steps
            if lo >= hi:
                continue
            # recompute the segment's states from its checkpoint
            X0, prev = self.checkpoints[lo]
            X = X0.clone()
            states = [X.clone()]
            for m in range(lo, hi):
                self._step(X, bz, m * dt, prev)
                if self.predictor == "extrapolate":
                    prev = states[-1]
                states.append(X.clone())
            if hi in self.checkpoints and not torch.equal(states[-1], self.checkpoints[hi][0]):
                raise RuntimeError("recomputed segment differs from its checkpoint")
            for m in range(hi - 1, lo - 1, -1):
                t = m * dt
                self._set_field(bz, t)
                x1, x0 = states[m + 1 - lo], states[m - lo]

                mu, it = ctx.adjoint_solve(x1, lam, hdt, self.adjoint_tol, precond=self.adjoint_precond)
                iters += it
                mu = mu.contiguous()
                f1, k1 = ctx.vjp_f(x1, mu, g, field=True, kappa=True, dsc=tmp)
                dsc += hdt * tmp
                f0, k0 = ctx.vjp_f(x0, mu, g, field=True, kappa=True, dsc=tmp)
                dsc += hdt * tmp
                lam = mu + hdt * g
                d_kappa += hdt * (k1 + k0)
                d_bz += hdt * (f1[2] + f0[2]) * self.profile(t)
        self.stats["adjoint_krylov_iters"] = iters
        # dL/dw at the interior nodes: sc = base * (1 - w)
        d_sc_ref = np.empty(n)
        d_sc_ref[self._perm] = dsc.cpu().numpy()
        full = self.device.idx.interior_to_full
        d_w = np.zeros(self.params.dim_x)
        d_w[full] = -d_sc_ref * self._base_sc[full]
        return d_kappa, d_bz, d_w


class TrapezoidalBz:
    """torch.autograd.Function: (kappa, Bz, w) -> B_z observable after the
    problem's steps, with the hand-written adjoint as its backward."""

    @staticmethod
    def apply(problem: AdjointProblem, kappa, bz, w):
        return _fn().apply(kappa, bz, w, problem)


_FN = None


def _fn():
    global _FN
    if _FN is not None:
        return _FN
    import torch

    class _TrapBz(torch.autograd.Function):
        @staticmethod
        def forward(ctx, kappa, bz, w, problem):
            ctx.problem = problem
            _, bzo = problem.forward(float(kappa), float(bz), w.detach().cpu().numpy())
            ctx.generation = problem.generation
            # each gradient goes back on its input's device and dtype
            ctx.like = [(v.device, v.dtype) if torch.is_tensor(v) else None for v in (kappa, bz, w)]
            return bzo

        @staticmethod
        def backward(ctx, g):
            p = ctx.problem
            if p.generation != ctx.generation:
                raise RuntimeError(
                    "the AdjointProblem ran forward again after this output was computed; its "
                    "checkpoints now belong to the later run: call backward before the next forward")
            dk, db, dw = p.backward(g.detach().to(torch.float64).contiguous())
            out = []
            for v, like in zip((dk, db, dw), ctx.like):
                out.append(None if like is None else
                           torch.as_tensor(v, dtype=torch.float64).to(device=like[0], dtype=like[1]))
            return out[0], out[1], out[2], None

    _FN = _TrapBz
    return _FN


def midplane_target_loss(problem: AdjointProblem, bz_obs, k_mid: Optional[int] = None,
                         amplitude: float = 0.2, width: float = 15.0):
    """L = mean over the interior plane k_mid of (B_z - T)^2 with the Gaussian
    bump T = amplitude * exp(-r^2 / (2 width^2)), r in cells from the film centre."""
    import torch
    p = problem.params
    nx, ny = p.Nx - 1, p.Ny - 1
    nz = p.Nz - 1 if p.Nz > 1 else 1
    k_mid = (p.Nz // 2) if k_mid is None else k_mid
    i, j = np.meshgrid(np.arange(1, p.Nx), np.arange(1, p.Ny), indexing="ij")
    r2 = (i - p.Nx / 2.0) ** 2 + (j - p.Ny / 2.0) ** 2
    T = amplitude * np.exp(-r2 / (2.0 * width ** 2))
    sel = ((i - 1) * ny + (j - 1)) * nz + (k_mid - 1 if p.Nz > 1 else 0)
    sel_t = torch.as_tensor(sel.ravel(), device=bz_obs.device)
    T_t = torch.as_tensor(T.ravel(), dtype=torch.float64, device=bz_obs.device)
    return ((bz_obs[sel_t] - T_t) ** 2).mean()
