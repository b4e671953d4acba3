"""Many films, one launch per kernel: the mixed-precision trapezoidal method
over a batch.

Each film keeps tdgl3d's own integrator semantics (integrators.py:92-206):
its applied field is evaluated at its own time, its dt is min(dt, t_stop - t)
and is divided by 10 on a failed step attempt (down to dt_min), and it stops
at t_stop.  The films advance in lockstep attempts: one
``tdgl_batch_trap_step`` per attempt for all films that have not finished.
"""

from __future__ import annotations

import hashlib
import json
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from .. import _native
from ..engine import engine_for


@dataclass
class BatchFilm:
    """One film of a batch: a Device and its initial state (reference layout)."""

    device: object
    x0: np.ndarray


@dataclass
class BatchResult:
    states: list                      # final state per film (reference layout, complex128)
    times: list                       # final time per film
    steps: list                       # accepted steps per film
    attempts: int                     # lockstep step attempts
    log: list = field(default_factory=list)  # per attempt: list of per-film stat dicts


class BatchRunner:
    """A tdgl_batch over Devices, with their FilmEngines kept alive."""

    def __init__(self, devices: Sequence):
        self.devices = list(devices)
        self.engines = [engine_for(d.params, d.idx, getattr(d, "material", None)) for d in self.devices]
        self.batch = _native.Batch([e.ctx for e in self.engines])

    def pack(self, x0s):
        return self.batch.pack([e.ctx.to_internal(x) for e, x in zip(self.engines, x0s)])

    def unpack(self, X, f: int):
        return self.engines[f].ctx.to_reference(self.batch.view(X, f)).cpu().numpy()

    def set_fields(self, times, t_stop) -> None:
        for d, e, t in zip(self.devices, self.engines, times):
            e.set_field(*d.applied_field.evaluate(t, t_stop))

    def tune_key(self, splits, shapes) -> str:
        """Cache key of a tuning: the GPU, the library build, the precision and
        every film's grid (Nx, Ny, Nz), with the candidates searched."""
        import torch
        dims = [[e.params.Nx, e.params.Ny, e.params.Nz] for e in self.engines]
        blob = json.dumps([torch.cuda.get_device_name(self.batch.device), _native.library_digest(),
                           "mixed", dims, [list(x) for x in splits], list(shapes)])
        return hashlib.sha256(blob.encode()).hexdigest()[:32]

    def _apply(self, shape, split) -> None:
        self.batch.set_tile(int(shape))
        self.batch.set_split(int(split[0]), int(split[1]))

    def tune(self, X=None, splits=((8, 2), (8, 4), (6, 2), (12, 2), (16, 4)), shapes=(0, 1, 2, 3),
             reps: int = 5, cache: Optional[str] = "default") -> dict:
        """Autotuner: pick the mixed-precision sweeps' tile shape (0 = 64x4,
        1 = 32x8, 2 = 16x16 nodes per plane, 3 = each film a tile of its own
        width, 16 to 64 nodes, with the fewest lanes per plane) and work split
        (the deepest and the shallowest z piece per block;
        tdgl_batch_set_sweep_split).

        Every shape with every split, and the current setting, is timed on the
        batch's own FP32 JVP sweep over all films; the fastest is kept (every
        setting gives the same bits).  The pick is cached in a JSON file
        (``cache``: a path, "default" = $NANOWIRE_TDGL_AMD_TUNE_CACHE or
        ~/.cache/nanowire_tdgl_amd/autotune.json, None = no cache) under
        :meth:`tune_key`; a later tuning of the same films on the same GPU reads
        it instead of timing.  Returns {"shape": s, "split": [max, min], "kz": [...]
        (the FP64 stencils' z depths, not tuned), "ms": best, "trials":
        {"<shape>:<g>x<l>": ms}, "cached": bool}."""
        b = self.batch
        F = b.nfilm
        kz = [b.get_kz(f) for f in range(F)]
        path = _cache_path(cache)
        key = self.tune_key(splits, shapes) if path else None
        if path and path.exists():
            try:
                hit = json.loads(path.read_text()).get(key)
            except (OSError, ValueError):
                hit = None
            if hit and "split" in hit:
                self._apply(hit["shape"], hit["split"])
                return dict(hit, kz=kz, cached=True)
        names = {0: "64x4", 1: "32x8", 2: "16x16", 3: "per-film"}
        best = (b.get_tile(), b.get_split())
        best_ms = b.time_stencil(reps)
        trials = {"default": best_ms}
        for sh in shapes:
            b.set_tile(sh)
            for g, l in splits:
                b.set_split(g, l)
                ms = b.time_stencil(reps)
                trials[f"{names[sh]}:{g}x{l}"] = ms
                if ms < best_ms:
                    best_ms, best = ms, (sh, (g, l))
        self._apply(*best)
        res = {"shape": int(best[0]), "split": [int(best[1][0]), int(best[1][1])], "ms": best_ms,
               "trials": trials}
        if path:
            _cache_store(path, key, res)
        return dict(res, kz=kz, cached=False)


def _cache_path(cache) -> Optional[Path]:
    if cache is None:
        return None
    if cache == "default":
        env = os.environ.get("NANOWIRE_TDGL_AMD_TUNE_CACHE")
        return Path(env) if env else Path.home() / ".cache" / "nanowire_tdgl_amd" / "autotune.json"
    return Path(cache)


def _cache_store(path: Path, key: str, res: dict) -> None:
    """Add one tuning to the JSON cache (written to a temporary file, then renamed)."""
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        data = json.loads(path.read_text()) if path.exists() else {}
    except (OSError, ValueError):
        data = {}
    data[key] = res
    tmp = path.with_suffix(path.suffix + f".{os.getpid()}.tmp")
    try:
        tmp.write_text(json.dumps(data))
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs the next run a tuning


def trapezoidal_batch(films: Sequence[BatchFilm], t_start: float, t_stop: float, dt: float, *,
                      newton_tol_f: float = 1e-3, newton_tol_dx: float = 1e-3,
                      newton_max_iter: int = 20, tol_gcr: float = 1e-4,
                      tgcr_max_iter: Optional[int] = None, mixed_tol_floor: float = 1e-6,
                      dt_min: float = 1e-8, max_attempts: Optional[int] = None,
                      tune: bool = False, runner: Optional[BatchRunner] = None,
                      keep_log: bool = False, precond="chebyshev",
                      predictor: str = "euler") -> BatchResult:
    """Advance every film from t_start to t_stop (mixed precision, lockstep).

    The Newton/TGCR parameters mean what they mean in
    :func:`nanowire_tdgl_amd.solvers.integrators.trapezoidal`; ``tune`` runs
    the stencil autotuner first; ``max_attempts`` bounds the number of
    lockstep attempts (None: until every film reaches t_stop); ``precond`` as
    in ``trapezoidal`` (each film takes its own automatic degree);
    ``predictor`` as in ``trapezoidal``, per film."""
    if predictor not in ("euler", "extrapolate"):
        raise ValueError("predictor must be 'euler' or 'extrapolate'")
    runner = runner or BatchRunner([f.device for f in films])
    F = len(films)
    X = runner.pack([f.x0 for f in films])
    if tune:
        runner.tune()
    opts = _native.solver_opts(newton_tol_f, newton_tol_dx, newton_max_iter, tol_gcr, 1e-4,
                               tgcr_max_iter, precision="mixed", mixed_tol_floor=mixed_tol_floor,
                               precond=precond)
    t = np.full(F, float(t_start))
    cur = np.full(F, float(dt))
    steps = np.zeros(F, dtype=np.int64)
    x_prev, dt_prev = None, np.zeros(F)
    attempts = 0
    log = []
    while True:
        active = t < t_stop - 1e-14
        if not active.any() or (max_attempts is not None and attempts >= max_attempts):
            break
        dta = np.where(active, np.minimum(cur, t_stop - t), dt)
        runner.set_fields(t, t_stop)
        x_in = X.clone() if predictor == "extrapolate" else None
        st = runner.batch.trap_step(X, dta, opts, active.astype(np.int32), x_prev=x_prev,
                                    dt_prev=dt_prev)
        attempts += 1
        if x_in is not None:  # a converged film's history becomes the step it just took
            for f in range(F):
                if active[f] and st[f].converged:
                    if x_prev is None:
                        x_prev = x_in.clone()
                    runner.batch.view(x_prev, f).copy_(runner.batch.view(x_in, f))
                    dt_prev[f] = dta[f]
        if keep_log:
            log.append([s.as_dict() for s in st])
        for f in range(F):
            if not active[f]:
                continue
            if st[f].converged:
                t[f] += dta[f]
                steps[f] += 1
                cur[f] = dt
            else:
                cur[f] /= 10.0
                if cur[f] < dt_min:
                    raise RuntimeError(f"film {f}: dt reduced below dt_min={dt_min:.1e} at t={t[f]:.6e}")
    states = [runner.unpack(X, f) for f in range(F)]
    return BatchResult(states, list(t), list(steps), attempts, log)
