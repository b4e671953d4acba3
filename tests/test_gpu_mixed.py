"""GPU: the mixed-precision trapezoidal path (analytic FP32 Jacobian-vector
product and FP32 Krylov vectors inside an FP64 Newton iteration).

* The FP32 JVP kernel is compared with the oracle's exact FP64 JVP evaluated
  at the same FP32-rounded inputs: <= 2e-6 of the largest entry (FP32 unit
  roundoff 6e-8 times the growth of a 7-point stencil with phase factors).
* Mixed steps are compared with FP64 steps of the same library.  The Newton
  residual is FP64 in both, so with a tight residual tolerance (||g||_inf <=
  1e-10) both converge to the same solution of the FP64 step equations
  (<= 1e-8 relative: 1e-10 times the conditioning of the step).  Much tighter
  tolerances are out of reach for either path: TGCR's absolute breakdown test
  ||Ap|| < 1e-14 (tgcr.py:141-142) ends the solve once the residual is that
  small.  At the
  reference's default Newton tolerances the two trajectories agree in |psi|^2,
  B_z and J_s to <= 1e-6 relative (of the largest entry).
"""

from __future__ import annotations

import numpy as np
import pytest

import specs as S
from helpers import device, oracle_film, rand_state, rel_err
from oracle import tdgl_oracle as O

pytestmark = pytest.mark.gpu


def _round32(v):
    v = np.asarray(v, np.complex128)
    return v.real.astype(np.float32).astype(np.float64) + 1j * v.imag.astype(np.float32)


def _engine(dev):
    from nanowire_tdgl_amd.engine import engine_for
    return engine_for(dev.params, dev.idx, dev.material)


@pytest.mark.parametrize("name", list(S.EVALF_CASES))
def test_jvp32_matches_exact_jvp(name):
    import torch
    case = S.EVALF_CASES[name]
    dev = device(case)
    f = oracle_film(case)
    eng = _engine(dev)
    eng.set_field(*case["field"])
    u = O.boundary_vectors(*case["field"], f.g, f.idx)
    n = dev.params.n_interior
    nphi = 3 if dev.params.Nz > 1 else 2
    x = _round32(rand_state(n, nphi, 21, 0.3, 0.2))
    p = _round32(rand_state(n, nphi, 22, 1.0, 1.0))
    ctx = eng.ctx
    gx, gp = ctx.to_internal(x).float(), ctx.to_internal(p).float()
    ap = torch.empty_like(gx)
    dt = 0.2
    ctx.jvp_trap32(gx, gp, dt, ap)
    got = ctx.to_reference(ap.double()).cpu().numpy()
    want = p - (dt / 2.0) * O.jvp_stencil(x, p, f.g, f.idx, u, f.material)
    err = rel_err(got, want)
    print(f"{name}: jvp32 relative error {err:.2e}")
    assert err < 2e-6


@pytest.mark.parametrize("name", ["c1_bz", "hole_bz", "aniso_bxyz", "periodic_xyz", "hole2d", "trilayer_hole"])
def test_jvp32_general_field_vectors(name):
    """The same applied field given as full-grid BoundaryVectors takes the
    sweeps' per-cell ghost rules (the general-field kernels) instead of the
    uniform-field masks: both agree with the oracle's exact JVP to 2e-6, and
    with each other to 1e-6 (the flux is the same float at every face cell),
    for the plain JVP and for one Chebyshev application."""
    import torch
    from nanowire_tdgl_amd.physics.applied_field import build_boundary_field_vectors
    case = S.EVALF_CASES[name]
    dev = device(case)
    f = oracle_film(case)
    eng = _engine(dev)
    ctx = eng.ctx
    u = O.boundary_vectors(*case["field"], f.g, f.idx)
    n = dev.params.n_interior
    nphi = 3 if dev.params.Nz > 1 else 2
    x = _round32(rand_state(n, nphi, 31, 0.3, 0.2))
    p = _round32(rand_state(n, nphi, 32, 1.0, 1.0))
    gx, gp = ctx.to_internal(x).float(), ctx.to_internal(p).float()
    dt = 0.2

    def sweep():
        ap = torch.empty_like(gx)
        ctx.jvp_trap32(gx, gp, dt, ap)
        w, z = torch.empty_like(gx), torch.empty_like(gx)
        ctx.precond_cheb32(gx, gp, dt, 5, w, z)
        return [ctx.to_reference(t.double()).cpu().numpy() for t in (ap, w, z)]

    ctx.set_field(*case["field"])
    uni = sweep()
    Bx, By, Bz = (np.array(v) for v in build_boundary_field_vectors(*case["field"], dev.params, dev.idx))
    ctx.set_field_vectors(Bx, By, Bz)
    gen = sweep()
    eng._field = None  # the engine's uniform-field cache no longer describes the context
    want = p - (dt / 2.0) * O.jvp_stencil(x, p, f.g, f.idx, u, f.material)
    errs = [rel_err(gen[0], want)] + [rel_err(a, b) for a, b in zip(gen, uni)]
    print(f"{name}: general-field jvp32 vs oracle {errs[0]:.2e}; vs uniform {errs[1:]}")
    assert errs[0] < 2e-6
    assert max(errs[1:]) < 1e-6


def test_jvp32_rejects_wrong_vectors():
    import torch
    dev = device(S.EVALF_CASES["c1_bz"])
    ctx = _engine(dev).ctx
    x = torch.zeros(ctx.len, dtype=torch.float64, device=ctx.device)
    with pytest.raises(ValueError):
        ctx.jvp_trap32(x, x, 0.1, x)


def _run(dev, X0, field, precision, steps, dt, predictor="euler", **kw):
    from nanowire_tdgl_amd import _native
    import nanowire_tdgl_amd as T
    eng = _engine(dev)
    af = T.AppliedField(**field)
    ctx = eng.ctx
    X = ctx.to_internal(X0)
    opts = _native.solver_opts(precision=precision, **kw)
    stats = []
    t = 0.0
    x_prev = None
    for _ in range(steps):
        eng.set_field(*af.evaluate(t, steps * dt))
        x_in = X.clone() if predictor == "extrapolate" else None
        st = ctx.trap_step(X, dt, opts, x_prev=x_prev, dt_prev=dt)
        assert st.converged == 1
        stats.append(st.as_dict())
        x_prev = x_in
        t += dt
    b, j = ctx.observables(X)
    psi = ctx.to_reference(X).cpu().numpy()[: dev.params.n_interior]
    n = dev.params.n_interior
    return np.abs(psi) ** 2, b.cpu().numpy()[2 * n:], j.cpu().numpy(), stats


def test_mixed_converges_to_the_fp64_solution():
    case = S.TRAP_HOLE
    dev = device(case)
    n = dev.params.n_interior
    X0 = rand_state(n, 3, 31, 0.05, 0.02)
    tight = dict(newton_tol_f=1e-10, newton_tol_dx=1e-6, tol_gcr=1e-4)
    a = _run(dev, X0, case["field"], "f64", 3, case["dt"], **tight)
    b = _run(dev, X0, case["field"], "mixed", 3, case["dt"], **tight)
    for k, name in enumerate(("|psi|^2", "B_z", "J_s")):
        err = rel_err(b[k], a[k])
        print(f"tight Newton: {name} relative difference {err:.2e}")
        assert err < 1e-8, name


@pytest.mark.parametrize("cname", ["TRAP_HOLE", "TRAP_C1"])
def test_mixed_trajectory_error_vs_fp64(cname):
    case = getattr(S, cname)
    dev = device(case)
    n = dev.params.n_interior
    rng = np.random.default_rng(case["seed"])
    X0 = np.concatenate([1.0 + case["noise"] * (rng.standard_normal(n) + 1j * rng.standard_normal(n)),
                         np.zeros(3 * n, np.complex128)])
    steps = 20
    a = _run(dev, X0, case["field"], "f64", steps, case["dt"], tol_gcr=case["tol_gcr"])
    b = _run(dev, X0, case["field"], "mixed", steps, case["dt"], tol_gcr=case["tol_gcr"])
    for k, name in enumerate(("|psi|^2", "B_z", "J_s")):
        err = rel_err(b[k], a[k])
        print(f"{cname}: {name} relative difference after {steps} steps {err:.2e}")
        assert err < 1e-6, name
    print("newton f64/mixed:", sum(s["newton_iters"] for s in a[3]),
          sum(s["newton_iters"] for s in b[3]),
          "tgcr:", sum(s["tgcr_iters_total"] for s in a[3]), sum(s["tgcr_iters_total"] for s in b[3]))


@pytest.mark.parametrize("predictor", ["euler", "extrapolate"])
def test_mixed_vs_converged_fp64(predictor):
    """At the reference's Newton tolerances (newton.py:84: 1e-3 on ||g||_inf and
    ||dx||_inf) the FP64 path's last iterate is a finite distance from the
    solution of the step equations; the mixed path, with its exact Jacobian, ends
    nearer to it.  Against an FP64 run whose Newton iteration is converged
    (||dx||_inf <= 1e-5), the mixed trajectory is within 1e-6 relative in |psi|^2,
    B_z and J_s (the headline film's settings on a 64x64x8 film with a 16x16 hole).
    The same holds with Newton started from the extrapolated last step
    (predictor="extrapolate"), which takes fewer Newton iterations."""
    case = dict(p=dict(Nx=64, Ny=64, Nz=8, hx=0.5, hy=0.5, hz=0.5, kappa=4.0),
                holes=[([(12.0, 12.0), (20.0, 12.0), (20.0, 20.0), (12.0, 20.0)], None)],
                field=dict(Bz=0.4), seed=1, noise=1e-3, dt=0.2, tol_gcr=1e-8)
    dev = device(case)
    n = dev.params.n_interior
    rng = np.random.default_rng(case["seed"])
    X0 = np.concatenate([1.0 + case["noise"] * (rng.standard_normal(n) + 1j * rng.standard_normal(n)),
                         np.zeros(3 * n, np.complex128)])
    steps, tol = 12, case["tol_gcr"]
    conv = _run(dev, X0, case["field"], "f64", steps, case["dt"], tol_gcr=tol, newton_tol_dx=1e-5)
    loose = _run(dev, X0, case["field"], "f64", steps, case["dt"], tol_gcr=tol)
    mixed = _run(dev, X0, case["field"], "mixed", steps, case["dt"], tol_gcr=tol, predictor=predictor)
    for k, name in enumerate(("|psi|^2", "B_z", "J_s")):
        em, ef = rel_err(mixed[k], conv[k]), rel_err(loose[k], conv[k])
        print(f"{name}: vs converged FP64, mixed ({predictor}) {em:.2e}, FP64 at the same tolerances {ef:.2e}")
        assert em < 1e-6, name
    nm, nf = sum(s["newton_iters"] for s in mixed[3]), sum(s["newton_iters"] for s in loose[3])
    print(f"Newton iterations: mixed ({predictor}) {nm}, FP64 {nf}")
    if predictor == "extrapolate":
        assert nm < nf
