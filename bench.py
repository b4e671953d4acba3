#!/usr/bin/env python3
"""Benchmark: trapezoidal TDGL steps on one MI355X.

Headline workload (the batched mixed-precision step): 64 uneven films from
seed 3 -- Nx, Ny uniform in [32, 160], Nz in [4, 24], h = 0.5, kappa ~ U[1, 5],
Bz ~ U[0.05, 0.8], a centred rectangular hole on the even films, psi = 1 +
1e-3 noise, phi = 0 -- advanced in lockstep attempts (one native call per
attempt for all films), dt = 0.2 with the integrator's per-film dt/10 retry,
TGCR tolerance 1e-4, Newton's defaults, the Chebyshev-preconditioned mixed
TGCR (FP32 JVP and Krylov vectors; FP64 Newton residual, Gram coefficients and
state update), tiles picked by the autotuner.  The films run as
`--batch-groups` (default 4) concurrent sub-batches balanced by nodes, each
its own batch on its own stream and host thread (trapezoidal_batch's
groups=); every film's arithmetic, and so every film's state, is the same bit
for bit as in one batch, whose figure is reported beside it.  A step
(`--steps`) is one lockstep attempt of every group; `value` is cell-steps/s:
interior nodes of the films whose attempt was accepted, summed over the timed
attempts, over the wall time, states resident in HBM.

One JSON line on stdout.  `roofline` is the dominant kernel class (the
Chebyshev sweeps), measured with HIP events on each launch over a second,
instrumented pass of the same attempts.  The line also carries the same
attempts looped over the plain FP64 step, and the per-film error of |psi|^2,
B_z and J_s against the oracle of this comparison: the plain FP64 GPU step
with Newton converged (||dx||_inf <= 1e-5; the solution of the same step
equations), with the FP64 loop at the same tolerances beside it.
`cpu_baseline` is the CPU oracle (faithful FP64 restatement of tdgl3d) timed
on whole steps of the batch's smallest films on this host, one thread and all
cores.  `batched.full_run` is the batch's whole configured run (every film's
300 steps, t = 0 to 60: the field is on until t = 40, then the films relax).

Secondary sections of the same line (each with its own roofline):
  fp64_plain -- the 128x128x16 production film with a 32x32 hole, plain FP64
                (the parity path), TGCR tol 1e-8;
  mixed      -- the same film and steps in mixed precision (and, beside it, FP64
                with its TGCR preconditioned by the same polynomial);
  large      -- 256x256x64 film with 1% insulating spheres (kappa 3, Bz 0.5,
                TGCR tol 1e-4): FP64 vs mixed steps and their error;
  adjoint    -- 96x96x12 film (kappa 2.5, Bz 0.3, mask w = 0.05 U(0,1) seed
                4): forward FP64 steps with checkpoints every 20, the
                hand-written adjoint of L = mean_midplane (B_z - T)^2, seconds
                per step of each, and central differences (eps = 1e-6) for
                kappa, Bz and random mask cells.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Off with --no-profile: no second, event-instrumented pass in any section.
PROFILE = True
# Off with --no-batch-full-run: the batch's whole configured run (300 steps per film).
FULL_RUN = True
# --batch-groups: the headline's concurrent sub-batches (1 = the films as one batch).
BATCH_GROUPS = 4
METRIC = "cell-steps/s mixed vs FP64 at 1e-6 rel err; achieved HBM GB/s; adjoint s/step"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_film():
    import nanowire_tdgl_amd as T
    p = T.SimulationParameters(Nx=128, Ny=128, Nz=16, hx=0.5, hy=0.5, hz=0.5, kappa=4.0)
    dev = T.Device(p, applied_field=T.AppliedField(Bz=0.4))
    dev.add_hole([(24.0, 24.0), (40.0, 24.0), (40.0, 40.0), (24.0, 40.0)])
    n = p.n_interior
    rng = np.random.default_rng(1)
    x0 = np.concatenate([1.0 + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)),
                         np.zeros(3 * n, complex)])
    return dev, x0


class Stepper:
    """The adaptive trapezoidal loop of the integrator, without history."""

    def __init__(self, dev, x0, dt=0.2, t_stop=200.0, tol_gcr=1e-8, precision="f64",
                 precond="chebyshev", predictor="euler", **newton):
        from nanowire_tdgl_amd import _native
        from nanowire_tdgl_amd.engine import engine_for
        self.dev = dev
        self.eng = engine_for(dev.params, dev.idx, dev.material)
        self.ctx = self.eng.ctx
        self.X = self.ctx.to_internal(x0)
        self.dt, self.cur, self.t, self.t_stop = dt, dt, 0.0, t_stop
        self.opts = _native.solver_opts(tol_gcr=tol_gcr, precision=precision, precond=precond,
                                        **newton)
        self.extrapolate = predictor == "extrapolate"
        self.x_prev, self.dt_prev = None, 0.0
        self.log = []

    def accepted_steps(self, k: int) -> None:
        done = 0
        while done < k:
            dta = min(self.cur, self.t_stop - self.t)
            self.eng.set_field(*self.dev.applied_field.evaluate(self.t, self.t_stop))
            x_in = self.X.clone() if self.extrapolate else None
            st = self.ctx.trap_step(self.X, dta, self.opts, x_prev=self.x_prev, dt_prev=self.dt_prev)
            self.log.append(st.as_dict())
            if st.converged:
                self.t += dta
                self.cur = self.dt
                done += 1
                if x_in is not None:
                    self.x_prev, self.dt_prev = x_in, dta
            else:
                self.cur /= 10.0
                if self.cur < 1e-8:
                    raise RuntimeError(f"dt reduced below dt_min at t={self.t:.6e}")

    def snapshot(self):
        xp = None if self.x_prev is None else self.x_prev.clone()
        return (self.X.clone(), self.t, self.cur, len(self.log), xp, self.dt_prev)

    def restore(self, snap) -> None:
        X, self.t, self.cur, nlog, xp, self.dt_prev = snap
        self.X.copy_(X)
        self.x_prev = None if xp is None else xp.clone()
        del self.log[nlog:]

    def observables(self):
        n = self.dev.params.n_interior
        b, j = self.ctx.observables(self.X)
        psi = self.ctx.to_reference(self.X).cpu().numpy()[:n]
        return np.abs(psi) ** 2, b.cpu().numpy()[2 * n:], j.cpu().numpy()


# Newton step tolerance of the converged FP64 runs that the errors are also
# measured against.  At the reference's 1e-3 (newton.py:84) the FP64 path's own
# last Newton iterate sits ~1e-6 (|psi|^2) from the solution of the step
# equations; FP64 runs at 1e-4 and 1e-5 agree to 1e-10.  Much below 1e-5 the
# Newton iteration ends in TGCR's absolute breakdown test (tgcr.py:141-142).
CONVERGED_NEWTON_DX = 1e-5


def rel_errors(a, b) -> dict:
    """max |a - b| / max |b| for |psi|^2, B_z, J_s (b = the FP64 run)."""
    out = {}
    for name, u, v in zip(("psi2", "bz", "js"), a, b):
        out[name] = float(np.max(np.abs(u - v)) / max(float(np.max(np.abs(v))), 1e-300))
    return out


def pmc_traffic(profile: str) -> dict:
    """HBM bytes per launch per kernel class from a committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by a copy of known size);
    empty if the summary is absent."""
    path = ROOT / "profiles" / profile
    if not path.exists():
        return {}
    for line in path.read_text().splitlines():
        if line.startswith("traffic-json:"):
            return json.loads(line.split(":", 1)[1])
    return {}


def with_traffic(roof, profile: str, section: str):
    """Attach the committed PMC traffic of the section's dominant kernel class."""
    if roof:
        tr = pmc_traffic(profile).get(section, {}).get(roof["kernel"])
        if tr:
            roof["traffic"] = tr
            roof["traffic_source"] = (f"profiles/{profile} (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE, "
                                      "calibrated on a 1 GiB copy)")
    return roof


def roofline_of(per_class: dict) -> dict:
    dom = max(per_class, key=lambda k: per_class[k]["ms"])
    c = per_class[dom]
    avg_ms = c["ms"] / c["launches"]
    bpl = c["bytes"] / c["launches"]
    ach = bpl / (avg_ms * 1e-3) / 1e9
    for k, v in per_class.items():
        if v["launches"]:
            v["avg_us"] = round(v["ms"] / v["launches"] * 1e3, 2)
            v["GBps"] = round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "avg_launch_us": round(avg_ms * 1e3, 2), "bytes_per_launch": bpl}


def timed_run(dev, x0, precision, steps, warmup, tol, profile=None, precond="chebyshev",
              predictor="euler", **opts):
    """Warm up, time `steps` accepted steps, then (optionally) re-run them
    instrumented from the same state.  Returns (Stepper, seconds, per_class, roofline)."""
    import torch
    s = Stepper(dev, x0, tol_gcr=tol, precision=precision, precond=precond, predictor=predictor, **opts)
    s.accepted_steps(warmup)
    torch.cuda.synchronize()
    snap = s.snapshot()
    t0 = time.perf_counter()
    s.accepted_steps(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    final = s.snapshot()
    per_class = roof = None
    if PROFILE if profile is None else profile:
        s.restore(snap)
        s.ctx.profile_enable(True)
        s.accepted_steps(steps)
        per_class = s.ctx.profile_read()
        s.ctx.profile_enable(False)
        roof = roofline_of(per_class)
        s.restore(final)
    s.timed = s.log[snap[3]:final[3]]
    return s, el, per_class, roof


def per_step(s, key, steps):
    return sum(r[key] for r in s.timed) / steps


def section_mixed(dev, x0, fp64_stepper, fp64_el, steps, warmup, tol):
    """Mixed precision with the Chebyshev-preconditioned TGCR (the default),
    and without the preconditioner, on the headline film and steps."""
    n = dev.params.n_interior
    s, el, pc, roof = timed_run(dev, x0, "mixed", steps, warmup, tol)
    roof = with_traffic(roof, "cheb_sweeps_mixed_sections.md", "mixed")
    v, v64 = n * steps / el, n * steps / fp64_el
    err = rel_errors(s.observables(), fp64_stepper.observables())
    log(f"[bench] mixed: {v:.4g} cell-steps/s ({v / v64:.2f}x FP64), err {err}")
    u, elu, _, _ = timed_run(dev, x0, "mixed", steps, warmup, tol, profile=False, precond="none")
    vu = n * steps / elu
    erru = rel_errors(u.observables(), fp64_stepper.observables())
    log(f"[bench] mixed, no preconditioner: {vu:.4g} cell-steps/s ({vu / v64:.2f}x FP64), err {erru}")
    # FP64 with its TGCR right-preconditioned by the same polynomial (an FP64
    # step of this build, not the reference's method)
    fpc, elp, _, _ = timed_run(dev, x0, "f64", steps, warmup, tol, profile=False, f64_precond="chebyshev")
    vp = n * steps / elp
    log(f"[bench] FP64, preconditioned TGCR: {vp:.4g} cell-steps/s ({vp / v64:.2f}x FP64 plain; "
        f"mixed {v / vp:.2f}x it)")
    # the FP64 solution of the same step equations, Newton converged (untimed)
    ref = Stepper(dev, x0, tol_gcr=tol, newton_tol_dx=CONVERGED_NEWTON_DX)
    ref.accepted_steps(warmup + steps)
    assert abs(ref.t - s.t) < 1e-9 and abs(fp64_stepper.t - s.t) < 1e-9
    obs = ref.observables()
    conv = {"mixed": rel_errors(s.observables(), obs),
            "mixed_unpreconditioned": rel_errors(u.observables(), obs),
            "fp64_same_tolerances": rel_errors(fp64_stepper.observables(), obs),
            "fp64_preconditioned": rel_errors(fpc.observables(), obs)}
    log(f"[bench] vs FP64 with Newton converged (dx <= {CONVERGED_NEWTON_DX:g}): {conv}")
    # Newton started from the extrapolated last step instead of X + dt f(X),
    # for the mixed method and, as its like-for-like baseline, for FP64
    ex = {}
    for prec in ("mixed", "f64"):
        e, ele, _, _ = timed_run(dev, x0, prec, steps, warmup, tol, profile=False, predictor="extrapolate")
        assert abs(e.t - s.t) < 1e-9
        ex[prec] = {"value": n * steps / ele, "newton_iters_per_step": per_step(e, "newton_iters", steps),
                    "tgcr_iters_per_step": per_step(e, "tgcr_iters_total", steps),
                    "rel_err_vs_fp64_converged": rel_errors(e.observables(), obs)}
    ex["mixed"]["speedup_vs_fp64_plain"] = ex["mixed"]["value"] / v64
    ex["mixed"]["speedup_vs_fp64_extrapolating"] = ex["mixed"]["value"] / ex["f64"]["value"]
    log(f"[bench] extrapolating predictor: mixed {ex['mixed']['value']:.4g} cell-steps/s "
        f"({ex['mixed']['speedup_vs_fp64_plain']:.2f}x FP64 plain, "
        f"{ex['mixed']['speedup_vs_fp64_extrapolating']:.2f}x FP64 extrapolating at {ex['f64']['value']:.4g}); "
        f"Newton/step {ex['mixed']['newton_iters_per_step']:.2f} (reference predictor "
        f"{per_step(s, 'newton_iters', steps):.2f}); err vs converged mixed {ex['mixed']['rel_err_vs_fp64_converged']}, "
        f"FP64 {ex['f64']['rel_err_vs_fp64_converged']}")
    return {"value": v, "unit": "cell-steps/s", "ms_per_step": el / steps * 1e3,
            "speedup_vs_fp64": v / v64, "precond": "chebyshev (automatic degree)",
            "rel_err_vs_fp64": err, "rel_err_vs_fp64_converged": conv,
            "roofline": roof, "kernel_classes": pc,
            "tgcr_iters_per_step": per_step(s, "tgcr_iters_total", steps),
            "stencil_sweeps_per_step": per_step(s, "jvp_evals", steps),
            "newton_iters_per_step": per_step(s, "newton_iters", steps),
            "extrapolating_predictor": ex,
            "unpreconditioned": {"value": vu, "speedup_vs_fp64": vu / v64, "rel_err_vs_fp64": erru,
                                 "tgcr_iters_per_step": per_step(u, "tgcr_iters_total", steps)},
            "fp64_preconditioned": {"value": vp, "speedup_vs_fp64": vp / v64, "mixed_speedup_over_it": v / vp,
                                    "tgcr_iters_per_step": per_step(fpc, "tgcr_iters_total", steps),
                                    "fd_jvps_per_step": per_step(fpc, "jvp_evals", steps)}}


def section_large(steps, warmup):
    from nanowire_tdgl_amd import workloads
    t0 = time.perf_counter()
    dev, x0 = workloads.film256_spheres()
    n = dev.params.n_interior
    log(f"[bench] large film built in {time.perf_counter() - t0:.1f} s, n_interior={n}, "
        f"insulating fraction {1 - dev.material.interior_sc_mask.mean():.4f}")
    tol = 1e-4
    a, ela, _, roofa = timed_run(dev, x0, "f64", steps, warmup, tol, profile=False)
    b, elb, pcb, roofb = timed_run(dev, x0, "mixed", steps, warmup, tol)
    roofb = with_traffic(roofb, "cheb_sweeps_mixed_sections.md", "large")
    c, elc, _, _ = timed_run(dev, x0, "mixed", steps, warmup, tol, profile=False, precond="none")
    va, vb, vc = n * steps / ela, n * steps / elb, n * steps / elc
    err = rel_errors(b.observables(), a.observables())
    log(f"[bench] large: FP64 {va:.4g}, mixed {vb:.4g} cell-steps/s ({vb / va:.2f}x), "
        f"mixed without preconditioner {vc:.4g} ({vc / va:.2f}x), err {err}")
    return {"workload": "256x256x64 film, h=0.5, kappa=3, Bz=0.5, 1% insulating spheres r in [2,4] "
                        "(seed 2), psi = (1 + 1e-3 noise) * sc_mask, dt=0.2, TGCR tol 1e-4",
            "n_interior": n, "steps": steps, "fp64_cell_steps_s": va, "mixed_cell_steps_s": vb,
            "speedup": vb / va, "rel_err_vs_fp64": err, "roofline_mixed": roofb,
            "kernel_classes_mixed": pcb,
            "mixed_unpreconditioned_cell_steps_s": vc, "speedup_unpreconditioned": vc / va,
            "tgcr_iters_per_step": {"f64": per_step(a, "tgcr_iters_total", steps),
                                    "mixed": per_step(b, "tgcr_iters_total", steps),
                                    "mixed_unpreconditioned": per_step(c, "tgcr_iters_total", steps)},
            "stencil_sweeps_per_step_mixed": per_step(b, "jvp_evals", steps)}


BATCH_DT = 0.2
BATCH_T_STOP = 300 * BATCH_DT  # the configured 300 steps: the field's on-window ends at 2/3 of it
BATCH_TOL = 1e-4
WORKLOAD_BATCH = ("64 films from seed 3: Nx,Ny in [32,160], Nz in [4,24], h=0.5, kappa~U[1,5], "
                  "Bz~U[0.05,0.8], holes on even films; psi = 1 + 1e-3 noise, phi = 0; dt=0.2 "
                  "(per-film dt/10 retry), TGCR tol 1e-4, Chebyshev-preconditioned mixed TGCR, "
                  "autotuned tiles")


def _quantiles(vals) -> dict:
    v = np.asarray(vals, dtype=np.float64)
    return {"median": float(np.median(v)), "p90": float(np.quantile(v, 0.9)), "max": float(v.max())}


class BatchAttempts:
    """Lockstep attempts over the batch with the integrator's per-film control
    flow (integrators.py:150-197): the field at the film's own time, dt/10 on
    a failed attempt, dt restored after an accepted one."""

    def __init__(self, runner, films, opts, predictor="euler"):
        self.runner, self.opts = runner, opts
        self.F = len(films)
        self.ns = np.array([f.device.params.n_interior for f in films], dtype=np.float64)
        self.X = runner.pack([f.x0 for f in films])
        self.t = np.zeros(self.F)
        self.cur = np.full(self.F, BATCH_DT)
        self.extrapolate = predictor == "extrapolate"
        self.x_prev, self.dt_prev = None, np.zeros(self.F)
        self.newton = 0
        self.work = np.zeros(self.F)  # per film: nodes x JVP and Chebyshev sweeps so far

    def snapshot(self):
        xp = None if self.x_prev is None else self.x_prev.clone()
        return (self.X.clone(), self.t.copy(), self.cur.copy(), xp, self.dt_prev.copy())

    def restore(self, snap) -> None:
        X, self.t, self.cur, xp, self.dt_prev = snap
        self.X.copy_(X)
        self.t, self.cur, self.dt_prev = self.t.copy(), self.cur.copy(), self.dt_prev.copy()
        self.x_prev = None if xp is None else xp.clone()

    def done(self) -> bool:
        return not (self.t < BATCH_T_STOP - 1e-14).any()

    def attempt(self) -> float:
        """One attempt for every film that has not reached t_stop (dt = min(dt,
        t_stop - t), integrators.py:151); returns the cell-steps it accepted."""
        b = self.runner.batch
        active = self.t < BATCH_T_STOP - 1e-14
        self.runner.set_fields(self.t, BATCH_T_STOP)
        dta = np.where(active, np.minimum(self.cur, BATCH_T_STOP - self.t), BATCH_DT)
        x_in = self.X.clone() if self.extrapolate else None
        st = b.trap_step(self.X, dta, self.opts, active.astype(np.int32), x_prev=self.x_prev,
                         dt_prev=self.dt_prev)
        ok = np.array([q.converged for q in st], dtype=bool) & active
        self.newton += sum(q.newton_iters for q, a in zip(st, active) if a)
        self.work += self.ns * np.array([q.jvp_evals if a else 0 for q, a in zip(st, active)])
        if x_in is not None:
            if self.x_prev is None:
                self.x_prev = x_in
            else:
                for f in np.flatnonzero(ok):
                    b.view(self.x_prev, f).copy_(b.view(x_in, f))
            self.dt_prev[ok] = dta[ok]
        self.t[ok] += dta[ok]
        self.cur[ok] = BATCH_DT
        self.cur[active & ~ok] /= 10.0
        if (self.cur < 1e-8).any():
            raise RuntimeError("dt reduced below dt_min in the batch")
        return float(self.ns[ok].sum())

    def run(self, k: int) -> float:
        return sum(self.attempt() for _ in range(k))


class GroupAttempts:
    """BatchAttempts over the films split into concurrent sub-batches
    (BatchGroups: each group its own tdgl_batch on its own
    stream, stepped from its own thread).  An attempt is one attempt of every
    group, so every film makes the same attempts, with the same bits, as in
    one batch."""

    def __init__(self, grp, films, opts, predictor="euler"):
        import torch
        self.grp = grp
        self.parts = []
        for k, r in enumerate(grp.runners):
            with torch.cuda.stream(grp.streams[k]):
                self.parts.append(BatchAttempts(r, [films[i] for i in grp.parts[k]], opts, predictor))
        torch.cuda.synchronize()

    @property
    def newton(self) -> int:
        return sum(b.newton for b in self.parts)

    def snapshot(self):
        return [b.snapshot() for b in self.parts]

    def done(self) -> bool:
        return all(b.done() for b in self.parts)

    def run(self, k: int) -> float:
        return sum(self.grp.run(lambda g, r: self.parts[g].run(k)))

    def run_to_end(self):
        """Every group until its films reach t_stop; returns (cell-steps, most attempts)."""
        def go(g, r):
            b, acc, n = self.parts[g], 0.0, 0
            while not b.done():
                acc += b.attempt()
                n += 1
            return acc, n
        res = self.grp.run(go)
        return sum(a for a, _ in res), max(n for _, n in res)


def section_batched(attempts, warmup, fp64=True, variants=True):
    """The headline: the 64-film batch in mixed precision, as BATCH_GROUPS
    concurrent sub-batches (the single batch's figure beside it)."""
    import torch
    from nanowire_tdgl_amd import _native, workloads
    from nanowire_tdgl_amd.solvers.batch import BatchGroups, BatchRunner
    films = workloads.batch_films()
    ns = np.array([f.device.params.n_interior for f in films])
    runner = BatchRunner([f.device for f in films])
    tune = runner.tune(reps=10)
    tune_info = {"picked": {"shape": {0: "64x4", 1: "32x8", 2: "16x16", 3: "per-film"}[tune["shape"]],
                            "split": tune["split"]}, "cached": tune.get("cached", False),
                 "jvp_ms_by_shape_and_split": {str(k): round(m, 4) for k, m in tune.get("trials", {}).items()}}
    log(f"[bench] batch autotune: {tune_info['picked']} (cached {tune_info['cached']})")
    opts = _native.solver_opts(tol_gcr=BATCH_TOL, precision="mixed")

    def timed(b: BatchAttempts):
        b.start = b.snapshot()
        b.run(warmup)
        torch.cuda.synchronize()
        b.work_warm = b.work.copy() if isinstance(b, BatchAttempts) else None
        snap = b.snapshot()
        n0 = b.newton
        w0 = time.monotonic_ns()  # the window on the clock kernel traces use
        c0 = time.perf_counter()
        acc = b.run(attempts)
        torch.cuda.synchronize()
        el_ = time.perf_counter() - c0
        b.window_ns = [w0, time.monotonic_ns()]
        b.newton_timed = (b.newton - n0) / (attempts * len(films))
        return snap, acc, el_


    mix = BatchAttempts(runner, films, opts)
    snap, acc, el1 = timed(mix)
    v1 = acc / el1
    log(f"[bench] batched mixed, one batch: {v1:.4g} cell-steps/s over {len(films)} films, "
        f"{attempts} attempts in {el1:.3f} s")
    grp, groups_info = None, None
    if BATCH_GROUPS > 1:
        # films alike in sweeps per attempt share a group (their sweeps stay
        # full), the groups balanced by that work: measured over the one
        # batch's warm-up attempts, before any timed attempt
        wk = mix.work_warm
        if wk.sum() > 0:
            grp = BatchGroups([f.device for f in films], BATCH_GROUPS, weights=wk, keys=wk / ns)
        else:  # no warm-up attempt to measure: node-balanced groups
            wk = ns.astype(np.float64)
            grp = BatchGroups([f.device for f in films], BATCH_GROUPS)
        gt = grp.tune()  # the groups timed together
        gm = GroupAttempts(grp, films, opts)
        _, accg, el = timed(gm)
        v = accg / el
        if accg != acc:
            raise RuntimeError("the grouped attempts accepted other steps than the single batch")
        groups_info = {"groups": len(grp), "films_per_group": [len(p) for p in grp.parts],
                       "nodes_per_group": [int(ns[p].sum()) for p in grp.parts],
                       "partition": ("films sorted by sweeps per attempt, cut into runs of equal sweep work "
                                     f"(both over the one batch's {warmup} warm-up attempts)" if warmup > 0
                                     else "balanced by nodes (no warm-up attempt to measure)"),
                       "work_share_per_group": [round(float(wk[p].sum() / wk.sum()), 3) for p in grp.parts],
                       "autotune": {"shape": {0: "64x4", 1: "32x8", 2: "16x16", 3: "per-film"}[gt["shape"]],
                                    "split": gt["split"],
                                    "jvp_ms_by_shape_and_split": {k: round(m, 4) for k, m in gt["trials"].items()}},
                       "one_batch_cell_steps_s": v1, "one_batch_seconds": el1,
                       "timed_window_monotonic_ns": gm.window_ns,
                       "one_batch_timed_window_monotonic_ns": mix.window_ns}
        log(f"[bench] batched mixed, {len(grp)} concurrent groups: {v:.4g} cell-steps/s "
            f"({attempts} attempts in {el:.3f} s; {v / v1:.3f}x one batch; tiles "
            f"{groups_info['autotune']['shape']}, split {gt['split']})")
    else:
        v, el = v1, el1
    pc = roof = None
    if PROFILE:  # instrumented repeat of the warm-up and timed attempts from the same state
        final = mix.snapshot()
        mix.restore(mix.start)
        runner.batch.profile_enable(True)
        mix.run(warmup + attempts)
        pc = runner.batch.profile_read()
        runner.batch.profile_enable(False)
        roof = with_traffic(roofline_of(pc), "cheb_sweeps_mixed_sections.md", "batched")
        mix.restore(final)
        log(f"[bench] batched dominant kernel {roof['kernel']}: {roof['avg_launch_us']:.1f} us/launch, "
            f"{roof['achieved']:.0f} GB/s ({roof['frac']:.3f} of peak), traffic {roof['traffic']}")
    # the whole attempt's algorithmic bytes (every kernel class, averaged over the instrumented
    # repeat's warm-up and timed attempts) over the timed wall time per attempt
    whole = (sum(c["bytes"] for c in pc.values()) / (warmup + attempts) / (el / attempts) / 1e9
             if pc else None)
    if not variants:
        out = {"workload": WORKLOAD_BATCH, "films": len(films), "n_interior_total": int(ns.sum()),
               "attempts": attempts, "warmup": warmup, "seconds": el, "value": v,
               "newton_iters_per_film_attempt": mix.newton_timed, "roofline": roof, "kernel_classes": pc,
               "whole_attempt_algorithmic_GBps": whole, "concurrent_groups": groups_info,
               "autotune": tune_info}
        return out
    full = None
    if FULL_RUN:  # the configured run: every film from t = 0 to its 300th step
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        if grp is not None:
            fr = GroupAttempts(grp, films, opts)
            accf, nat = fr.run_to_end()
        else:
            fr = BatchAttempts(runner, films, opts)
            accf, nat = 0.0, 0
            while not fr.done():
                accf += fr.attempt()
                nat += 1
        torch.cuda.synchronize()
        elf = time.perf_counter() - c0
        full = {"steps_per_film": int(round(BATCH_T_STOP / BATCH_DT)), "attempts": nat, "seconds": elf,
                "groups": len(grp) if grp is not None else 1, "cell_steps_s": accf / elf,
                "newton_iters_per_film_step": fr.newton / (len(films) * BATCH_T_STOP / BATCH_DT)}
        log(f"[bench] batched mixed, whole configured run: {full['cell_steps_s']:.4g} cell-steps/s "
            f"({nat} attempts in {elf:.2f} s)")
    # the same attempts without the preconditioner, and with the extrapolating predictor
    unp = BatchAttempts(runner, films, _native.solver_opts(tol_gcr=BATCH_TOL, precision="mixed",
                                                           precond="none"))
    _, accu, elu = timed(unp)
    vu = accu / elu
    log(f"[bench] batched mixed, no preconditioner: {vu:.4g} cell-steps/s")
    ex = (GroupAttempts(grp, films, opts, predictor="extrapolate") if grp is not None else
          BatchAttempts(runner, films, opts, predictor="extrapolate"))
    ex.run(1)  # every film gets a last step to extrapolate from
    _, acce, ele = timed(ex)
    ve = acce / ele
    log(f"[bench] batched mixed, extrapolating predictor: {ve:.4g} cell-steps/s")
    out = {"workload": WORKLOAD_BATCH,
           "films": len(films), "n_interior_total": int(ns.sum()), "attempts": attempts,
           "warmup": warmup, "seconds": el, "value": v,
           "newton_iters_per_film_attempt": mix.newton_timed,
           "roofline": roof, "kernel_classes": pc, "whole_attempt_algorithmic_GBps": whole,
           "concurrent_groups": groups_info,
           "full_run": full,
           "unpreconditioned_cell_steps_s": vu,
           "extrapolating_cell_steps_s": ve,
           "extrapolating_newton_iters_per_film_attempt": ex.newton_timed,
           "autotune": tune_info}
    if not fp64:
        return out
    # baseline: the plain FP64 step looped over the films, same steps from the same start
    steppers = [Stepper(f.device, f.x0, t_stop=BATCH_T_STOP, tol_gcr=BATCH_TOL) for f in films]
    for s_ in steppers:
        s_.accepted_steps(warmup)
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    for s_ in steppers:
        s_.accepted_steps(attempts)
    torch.cuda.synchronize()
    el64 = time.perf_counter() - c0
    v64 = float(ns.sum()) * attempts / el64
    log(f"[bench] FP64 loop: {v64:.4g} cell-steps/s ({el64:.2f} s); batched/loop = {v / v64:.2f}")
    # the same loop with the FP64 TGCR preconditioned by the mixed method's polynomial (an FP64
    # step of this build): what the batch gains from precision and batching alone
    psteppers = [Stepper(f.device, f.x0, t_stop=BATCH_T_STOP, tol_gcr=BATCH_TOL, f64_precond="chebyshev")
                 for f in films]
    for s_ in psteppers:
        s_.accepted_steps(warmup)
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    for s_ in psteppers:
        s_.accepted_steps(attempts)
    torch.cuda.synchronize()
    v64p = float(ns.sum()) * attempts / (time.perf_counter() - c0)
    log(f"[bench] FP64 loop, preconditioned TGCR: {v64p:.4g} cell-steps/s; batched/it = {v / v64p:.2f}")
    # per-film error against the oracle of the comparison (FP64, Newton converged)
    # and against the FP64 loop at the same tolerances, films at the same time
    errs = {"mixed_vs_fp64_converged": [], "fp64_same_tol_vs_fp64_converged": [],
            "mixed_vs_fp64_same_tol": []}
    compared = 0
    for f, s_ in enumerate(steppers):
        eng = runner.engines[f]
        Xf = runner.batch.view(mix.X, f)
        ref = Stepper(films[f].device, films[f].x0, t_stop=BATCH_T_STOP, tol_gcr=BATCH_TOL,
                      newton_tol_dx=CONVERGED_NEWTON_DX)
        ref.accepted_steps(warmup + attempts)
        if abs(mix.t[f] - ref.t) > 1e-9 or abs(s_.t - ref.t) > 1e-9:
            continue
        compared += 1
        n = ns[f]
        b, j = eng.ctx.observables(Xf)
        psi = eng.ctx.to_reference(Xf).cpu().numpy()[:n]
        mine = (np.abs(psi) ** 2, b.cpu().numpy()[2 * n:], j.cpu().numpy())
        obs, same = ref.observables(), s_.observables()
        errs["mixed_vs_fp64_converged"].append(rel_errors(mine, obs))
        errs["fp64_same_tol_vs_fp64_converged"].append(rel_errors(same, obs))
        errs["mixed_vs_fp64_same_tol"].append(rel_errors(mine, same))
    dist = {k: {q: _quantiles([e[q] for e in v_]) for q in ("psi2", "bz", "js")}
            for k, v_ in errs.items() if v_}
    log(f"[bench] batched error distribution over {compared} films: {dist}")
    out.update({"fp64_loop_cell_steps_s": v64, "speedup_vs_fp64_loop": v / v64,
                "fp64_preconditioned_loop_cell_steps_s": v64p,
                "speedup_vs_fp64_preconditioned_loop": v / v64p,
                "speedup_extrapolating_vs_fp64_loop": ve / v64,
                "speedup_unpreconditioned_vs_fp64_loop": vu / v64,
                "films_compared": compared,
                "error_oracle": ("plain FP64 GPU step with Newton converged (||dx||_inf <= 1e-5), "
                                 "same films, steps and fields"),
                "rel_err": dist,
                "rel_err_psi2_per_film": [float(f"{e['psi2']:.3g}") for e in errs["mixed_vs_fp64_converged"]]})
    return out


def cpu_baseline(budget: float, workers: int) -> dict:
    """The oracle on whole steps of the batch's smallest films: one thread,
    then `workers` processes of one thread each."""
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    runs = {}
    for w in (1, workers):
        out = subprocess.run([sys.executable, "-m", "oracle.cpu_sample", "--budget", str(budget),
                              "--workers", str(w)], cwd=str(ROOT), env=env, capture_output=True,
                             text=True, timeout=1200)
        if out.returncode != 0:
            raise RuntimeError(out.stderr[-2000:])
        runs[w] = json.loads(out.stdout.strip().splitlines()[-1])
        log(f"[bench] CPU oracle, {w} worker(s): {runs[w]['value']:.4g} cell-steps/s "
            f"({runs[w]['films']} film steps in {runs[w]['elapsed_s']:.1f} s)")
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    one, allc = runs[1], runs[workers]
    return {"value": allc["value"], "unit": "cell-steps/s", "cores": workers, "kind": "port",
            "cpu_model": model, "cpu_count": os.cpu_count(),
            "sample": (f"oracle (CSR eval_f rebuilt per call, FD-JVP TGCR with MGS, Newton, dt/10 "
                       f"retry) on whole first steps (dt=0.2, field on, TGCR tol 1e-4) of the "
                       f"batch's films, smallest first, for a {budget:.0f} s budget: "
                       f"{allc['films']} film steps ({allc['cell_steps']} cell-steps) in "
                       f"{allc['elapsed_s']:.1f} s on {workers} one-thread processes"),
            "one_thread": {"value": one["value"], "films": one["films"],
                           "cell_steps": one["cell_steps"], "seconds": one["elapsed_s"]},
            "all_cores": {"value": allc["value"], "workers": workers, "films": allc["films"],
                          "cell_steps": allc["cell_steps"], "seconds": allc["elapsed_s"]}}


def section_adjoint(steps, every, fd=True, predictor=None, mask_cells=20):
    import torch
    import nanowire_tdgl_amd as T
    from nanowire_tdgl_amd.adjoint import AdjointProblem, TrapezoidalBz, midplane_target_loss
    p = T.SimulationParameters(Nx=96, Ny=96, Nz=12, hx=0.5, hy=0.5, hz=0.5, kappa=2.5)
    n = p.n_interior
    rng = np.random.default_rng(4)
    x0 = np.concatenate([1.0 + 0.01 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)),
                         np.zeros(3 * n, complex)])
    w0 = 0.05 * np.random.default_rng(4).uniform(size=p.dim_x)
    prob = AdjointProblem(p, x0, dt=0.2, steps=steps, checkpoint_every=every)
    if predictor is not None:
        prob.predictor = predictor
    # untimed warm-up on the same context: two steps forward and back load the
    # kernels and allocate the work buffers the timed run uses
    prob.steps = 2
    wk = torch.tensor(2.5, dtype=torch.float64, requires_grad=True)
    wb = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    ww = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    midplane_target_loss(prob, TrapezoidalBz.apply(prob, wk, wb, ww)).backward()
    prob.steps = steps
    prob.stats = {}
    kappa = torch.tensor(2.5, dtype=torch.float64, requires_grad=True)
    bz = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    loss = midplane_target_loss(prob, TrapezoidalBz.apply(prob, kappa, bz, w))
    torch.cuda.synchronize()
    c1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    c2 = time.perf_counter()
    out = {"workload": "96x96x12 film, h=0.5, kappa=2.5, Bz=0.3 (t_on_fraction 1), w = 0.05 U(0,1) "
                       f"(seed 4), psi = 1 + 0.01 noise (seed 4); {steps} FP64 steps dt=0.2, "
                       f"checkpoints every {every}; L = mean over k=6 of (B_z - T)^2, T = 0.2 "
                       "exp(-r^2/(2 15^2))",
           "n_interior": n, "steps": steps, "loss": float(loss.detach()),
           "forward_s_per_step": (c1 - c0) / steps, "adjoint_s_per_step": (c2 - c1) / steps,
           "grad": {"kappa": kappa.grad.item(), "Bz": bz.grad.item(),
                    "w_norm": float(w.grad.norm())},
           "predictor": prob.predictor,
           "newton_iters": prob.stats.get("newton"), "adjoint_krylov_iters": prob.stats.get("adjoint_krylov_iters")}
    log(f"[bench] adjoint: forward {out['forward_s_per_step']:.3f} s/step, adjoint "
        f"{out['adjoint_s_per_step']:.3f} s/step, grads {out['grad']}")
    if fd:
        def L(k, b):
            _, bzo = prob.forward(k, b, w0)
            return float(midplane_target_loss(prob, bzo))
        eps = 1e-6
        fdk = (L(2.5 + eps, 0.3) - L(2.5 - eps, 0.3)) / (2 * eps)
        fdb = (L(2.5, 0.3 + eps) - L(2.5, 0.3 - eps)) / (2 * eps)
        out["fd"] = {"eps": eps, "kappa": fdk, "Bz": fdb,
                     "rel_diff_kappa": abs(fdk - out["grad"]["kappa"]) / abs(fdk),
                     "rel_diff_Bz": abs(fdb - out["grad"]["Bz"]) / abs(fdb)}
        log(f"[bench] adjoint vs central differences: {out['fd']}")
        # random interior mask cells (seed 4), the same eps
        gw = w.grad.numpy()
        full = prob.device.idx.interior_to_full
        cells = np.random.default_rng(4).choice(full, size=20, replace=False)[:mask_cells]

        def Lw(wv):
            _, bzo = prob.forward(2.5, 0.3, wv)
            return float(midplane_target_loss(prob, bzo))

        diffs = []
        for c, m in enumerate(cells):
            wp, wm = w0.copy(), w0.copy()
            wp[m] += eps
            wm[m] -= eps
            diffs.append(abs((Lw(wp) - Lw(wm)) / (2 * eps) - gw[m]))
            if (c + 1) % 5 == 0:
                log(f"[bench] mask cells: {c + 1} of {len(cells)} done")
        if diffs:
            scale = float(np.max(np.abs(gw)))
            out["fd"]["mask_cells"] = {"cells": len(cells), "max_abs_diff": float(max(diffs)),
                                       "grad_scale": scale, "max_diff_over_scale": float(max(diffs)) / scale}
            log(f"[bench] adjoint vs central differences on {len(cells)} mask cells: max |diff| "
                f"{max(diffs):.3e} ({max(diffs) / scale:.2e} of max |dL/dw|)")
    return out


def whole_run(total: int = 1000, chunk: int = 100) -> None:
    """The production film's configured run (1000 steps, t = 0..200) for the
    FP64 plain, mixed, and mixed extrapolating methods, timed in chunks;
    prints the rate of every chunk and of the whole run."""
    import torch
    dev, x0 = build_film()
    n = dev.params.n_interior
    for prec, pred in (("f64", "euler"), ("mixed", "euler"), ("mixed", "extrapolate")):
        s = Stepper(dev, x0, precision=prec, predictor=pred)
        tot = 0.0
        for done in range(chunk, total + 1, chunk):
            torch.cuda.synchronize()
            c0 = time.perf_counter()
            s.accepted_steps(chunk)
            torch.cuda.synchronize()
            el = time.perf_counter() - c0
            tot += el
            nit = sum(r["newton_iters"] for r in s.log[-chunk:]) / chunk
            log(f"[bench] {prec} {pred}: steps {done}, t = {s.t:.1f}: {n * chunk / el:.4g} cell-steps/s, "
                f"{nit:.2f} Newton iterations per step")
        log(f"[bench] {prec} {pred}: {total} steps in {tot:.1f} s = {n * total / tot:.4g} cell-steps/s")


def calibrate_copy(mib: int, reps: int) -> None:
    """Copy a resident `mib` MiB buffer `reps` times (one copy kernel per
    rep): the reference point that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes (each copy reads and writes exactly mib MiB)."""
    import torch
    a = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_copy": mib * (1 << 20),
                      "bytes_written_per_copy": mib * (1 << 20), "copies": reps}), flush=True)


def section_fp64_plain(dev, x0, steps, warmup):
    """The plain FP64 method (the parity path) on the 128x128x16 production film."""
    n = dev.params.n_interior
    s, el, per_class, roof = timed_run(dev, x0, "f64", steps, warmup, 1e-8)
    timed = s.timed
    value = n * steps / el
    log(f"[bench] FP64 plain, 128x128x16: {el:.3f} s for {steps} steps ({len(timed)} attempts): "
        f"{value:.4g} cell-steps/s")
    if roof:
        tr = pmc_traffic("pmc_traffic_film128_hole.md").get(roof["kernel"])
        if tr:
            roof["traffic"] = tr
            roof["traffic_source"] = ("profiles/pmc_traffic_film128_hole.md (rocprofv3 FETCH_SIZE x 2.0 + "
                                      "WRITE_SIZE, calibrated on a 1 GiB copy)")
        log(f"[bench] FP64 dominant kernel {roof['kernel']}: {roof['avg_launch_us']:.1f} us/launch, "
            f"{roof['achieved']:.0f} GB/s, traffic {roof['traffic']}")
    step_bytes = sum(r["bytes_moved"] for r in timed)
    out = {"workload": "plain FP64 trapezoidal step, 128x128x16 film, h=0.5, kappa=4, Bz=0.4, "
                       "32x32 hole, psi = 1 + 1e-3 noise (seed 1), dt=0.2, TGCR tol 1e-8",
           "n_interior": n, "steps": steps, "warmup": warmup, "value": value,
           "ms_per_step": el / steps * 1e3, "roofline": roof, "kernel_classes": per_class,
           "attempts": len(timed),
           "newton_iters_per_step": sum(r["newton_iters"] for r in timed) / steps,
           "tgcr_iters_per_step": sum(r["tgcr_iters_total"] for r in timed) / steps,
           "whole_step_algorithmic_GBps": step_bytes / el / 1e9}
    return out, s, el


def main() -> None:
    global PROFILE, FULL_RUN, BATCH_GROUPS
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel used to calibrate the PMC byte counters")
    ap.add_argument("--whole-run", action="store_true",
                    help="only time the production film's whole 1000-step run per method, in chunks")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed lockstep attempts of the batch")
    ap.add_argument("--warmup", type=int, default=5, help="untimed attempts before them")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-workers", type=int, default=min(os.cpu_count() or 1, 16),
                    help="processes of the all-cores CPU figure (default: all cores, at most 16)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-batch-full-run", action="store_true",
                    help="skip the batch's whole configured run (300 steps per film)")
    ap.add_argument("--batch-groups", type=int, default=BATCH_GROUPS,
                    help="concurrent sub-batches of the headline (1: the films as one batch)")
    ap.add_argument("--no-batch-fp64", action="store_true",
                    help="skip the FP64 loop baseline and the error report of the batch")
    ap.add_argument("--no-batch-variants", action="store_true",
                    help="skip the batch's runs without the preconditioner and with the "
                         "extrapolating predictor (the profiled command of the headline)")
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-mask-cells", type=int, default=6)
    ap.add_argument("--sections", default="fp64_plain,mixed,large,adjoint",
                    help="comma list of secondary sections to run (empty: headline only)")
    ap.add_argument("--sections-only", action="store_true",
                    help="profiling aid: run only the secondary sections (no headline value)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, 5)
        return
    if args.whole_run:
        whole_run()
        return
    PROFILE = not args.no_profile
    FULL_RUN = not args.no_batch_full_run
    BATCH_GROUPS = max(1, args.batch_groups)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU split of "
              "this path is a design note, not a measured run", file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401
    if args.sections_only:
        head = None
    else:
        head = section_batched(args.steps, args.warmup, fp64=not args.no_batch_fp64,
                               variants=not args.no_batch_variants)
    sections = [x for x in args.sections.split(",") if x]
    extra = {}
    if "fp64_plain" in sections or "mixed" in sections:
        dev, x0 = build_film()
        fp, s64, el64 = section_fp64_plain(dev, x0, args.fp64_steps, 2)
        if "fp64_plain" in sections:
            extra["fp64_plain"] = fp
        if "mixed" in sections:
            extra["mixed"] = section_mixed(dev, x0, s64, el64, args.fp64_steps, 2, 1e-8)
    if "large" in sections:
        extra["large"] = section_large(args.large_steps, 1)
    if "adjoint" in sections:
        extra["adjoint"] = section_adjoint(args.adjoint_steps, 20, mask_cells=args.adjoint_mask_cells)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU oracle sample ({args.cpu_budget:.0f} s budget per run)")
        cpu = cpu_baseline(args.cpu_budget, args.cpu_workers)

    if head is None:
        print(json.dumps({"sections_only": True, **extra}), flush=True)
        return
    value = head.pop("value")
    el = head.pop("seconds")
    roof = head.pop("roofline")
    res = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "mixed",
        "data": "synthetic",
        "config": {"workload": head.pop("workload"), "films": head.pop("films"),
                   "n_interior_total": head.pop("n_interior_total"), "precision": "mixed",
                   "step": "one lockstep trapezoidal attempt over all films"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        res["speedup_vs_cpu_all_cores"] = value / cpu["value"]
    res["batched"] = head
    res.update(extra)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
