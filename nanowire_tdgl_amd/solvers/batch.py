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


# The piece splits (deepest / shallowest z piece per block) the autotuners try.
SPLITS = ((8, 2), (8, 4), (6, 2), (10, 4), (12, 2), (12, 4), (16, 4))
SHAPE_NAMES = {0: "64x4", 1: "32x8", 2: "16x16", 3: "per-film"}


class BatchRunner:
    """A tdgl_batch over Devices, with their FilmEngines kept alive; ``stream``
    (a torch stream) is the one the batch's steps are ordered on (default: the
    stream current when the first film's engine was made)."""

    def __init__(self, devices: Sequence, stream=None):
        self.devices = list(devices)
        self.engines = [engine_for(d.params, d.idx, getattr(d, "material", None)) for d in self.devices]
        self.batch = _native.Batch([e.ctx for e in self.engines], stream=stream)

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

    def tune(self, X=None, splits=SPLITS, shapes=(0, 1, 2, 3),
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
        names = SHAPE_NAMES
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


def partition_films(sizes: Sequence[float], groups: int, keys: Optional[Sequence[float]] = None) -> list:
    """Split films (their weights ``sizes``: node counts, or measured work)
    into ``groups`` groups of near-equal total weight.  Without ``keys``: the
    heaviest film first, each film to the group with the least weight so far
    (ties to the lowest group).  With ``keys`` (one per film, such as a film's
    sweeps per attempt): the films sorted by key are cut into contiguous runs
    of near-equal weight, each film to the run its weight's midpoint falls in,
    so films alike in key share a group.  Returns the groups as ascending lists
    of film indices; no group is empty when groups <= films."""
    if groups < 1:
        raise ValueError("groups must be >= 1")
    groups = min(groups, len(sizes))
    if keys is not None:
        if len(keys) != len(sizes):
            raise ValueError("one key per film")
        order = sorted(range(len(sizes)), key=lambda i: (float(keys[i]), i))
        tot = float(sum(float(w) for w in sizes)) or 1.0
        parts = [[] for _ in range(groups)]
        acc, k = 0.0, 0
        for pos, i in enumerate(order):
            want = int((acc + 0.5 * float(sizes[i])) * groups / tot)
            # never skip a group, and leave every later group a film
            k = min(k + (1 if parts[k] else 0), max(k, want, groups - (len(order) - pos)), groups - 1)
            parts[k].append(i)
            acc += float(sizes[i])
        return [sorted(p) for p in parts]
    load = [0.0] * groups
    parts: list = [[] for _ in range(groups)]
    for i in sorted(range(len(sizes)), key=lambda i: (-sizes[i], i)):
        k = min(range(groups), key=lambda k: (load[k], k))
        parts[k].append(i)
        load[k] += float(sizes[i])
    return [sorted(p) for p in parts]


class BatchGroups:
    """The films of a batch split into G sub-batches that run concurrently.

    Each group (:func:`partition_films`: balanced by nodes, or by ``weights``
    such as each film's measured work; with ``keys``, such as each film's
    sweeps per attempt, films alike in key share a group, so a group's films
    stop their Chebyshev sweeps and TGCR iterations at similar counts and its
    sweeps stay full) is a :class:`BatchRunner` of its own,
    stepped from a host thread of its own (the library's calls release the
    GIL), and the group's host-side torch work runs on its batch's own stream.
    One group's host work -- the Newton control flow, the converged-film
    bookkeeping, the readbacks of a lockstep attempt -- overlaps the other
    groups' kernels, and their kernels share the GPU.  A film's arithmetic does
    not depend on the other films of its batch, so every film's result is the
    same bit for bit whatever the grouping.

    Each group needs a hardware queue of its own: HIP maps a process's streams
    onto GPU_MAX_HW_QUEUES queues (default 4) and shares them once they run
    out, and two groups on one queue run one after the other.  With
    ``own_queues`` (the default) every group's batch moves to a stream with a
    queue of its own (tdgl_batch_own_queue).  Those streams are blocking: while
    the groups run, no other thread may use the legacy (default) stream, or a
    group's graph capture fails.  Without ``own_queues`` set GPU_MAX_HW_QUEUES
    to at least G + 2 before the process first uses the GPU (a warning says so
    otherwise).  On the 64-film batch, 2 to 4 groups gain 11-17 % over one
    batch; 5 and more lose."""

    def __init__(self, devices: Sequence, groups: int, weights: Optional[Sequence[float]] = None,
                 own_queues: bool = True, keys: Optional[Sequence[float]] = None):
        devices = list(devices)
        sizes = [d.params.n_interior for d in devices] if weights is None else list(weights)
        if len(sizes) != len(devices):
            raise ValueError("one weight per film")
        self.parts = partition_films(sizes, groups, keys)
        hwq = os.environ.get("GPU_MAX_HW_QUEUES", "")
        if not own_queues and len(self.parts) > 1 and int(hwq if hwq.isdigit() else 4) < len(self.parts) + 2:
            import warnings
            warnings.warn(f"{len(self.parts)} concurrent groups with GPU_MAX_HW_QUEUES={hwq or 4}: groups "
                          "may share a hardware queue and run one after the other; set it to at least "
                          f"{len(self.parts) + 2} before the process first uses the GPU", RuntimeWarning,
                          stacklevel=2)
        self.runners, self.streams = [], []
        for part in self.parts:
            r = BatchRunner([devices[i] for i in part])
            if own_queues and len(self.parts) > 1:
                try:
                    r.batch.own_queue()
                except _native.TDGLError as e:  # a runtime without CU-masked streams: shared queues
                    import warnings
                    warnings.warn(f"no stream with a queue of its own ({e}); the group shares HIP's "
                                  "hardware queues", RuntimeWarning, stacklevel=2)
            st = r.batch.work_stream()
            r.batch.set_stream(st)
            self.runners.append(r)
            self.streams.append(st)

    def __len__(self) -> int:
        return len(self.parts)

    # Grouped runs try pieces of at least 8 planes.  The least depth acts only
    # on sweeps over a few films (a full sweep's pieces are as deep as its work
    # per block slot), which the tuner's full JVP sweeps do not contain; in a
    # grouped attempt those sweeps' tails are covered by the other groups'
    # sweeps, so shallower pieces there only add per-block set-up.  Measured on
    # the 64-film batch as 4 groups: 211.5-212.7 M cell-steps/s with a least
    # depth of 4 planes, 216.7-218.2 M with 8 to 16.
    GROUP_SPLITS = ((16, 8), (24, 8), (32, 8), (32, 12), (64, 16))

    def tune(self, splits=GROUP_SPLITS, shapes=(3,), reps: int = 40, repeats: int = 5) -> dict:
        """The autotuner for groups that run together: every group takes each
        tile shape and piece split, their FP32 JVP sweeps over all their films
        are timed at once (the groups start together at a barrier; the slowest
        group's event time, the median of ``repeats``: the settings differ by a
        few per cent, and a best-of count picks whichever setting had one
        lucky overlap), and the fastest setting is kept for every group.  A group timed alone wants shallower pieces
        (more blocks for its share of the GPU) than it does beside the others,
        whose sweeps fill its start and tail.  Returns {"shape", "split", "ms",
        "trials"}."""
        import threading
        best, trials = None, {}
        for sh in shapes:
            for g, l in splits:
                for r in self.runners:
                    r.batch.set_tile(int(sh))
                    r.batch.set_split(int(g), int(l))
                runs = []
                for _ in range(repeats):
                    gate = threading.Barrier(len(self.runners))

                    def timed(k, r):
                        gate.wait()
                        return r.batch.time_stencil(reps)

                    runs.append(max(self.run(timed)))
                ms = float(np.median(runs))
                trials[f"{SHAPE_NAMES[sh]}:{g}x{l}"] = ms
                if best is None or ms < best[0]:
                    best = (ms, int(sh), (int(g), int(l)))
        for r in self.runners:
            r._apply(best[1], best[2])
        return {"shape": best[1], "split": list(best[2]), "ms": best[0], "trials": trials}

    def run(self, fn) -> list:
        """``fn(k, runner)`` for every group k at once, each in its own thread
        with the group's stream current; returns the results in group order
        once every group has finished (the first exception is raised again)."""
        import threading
        import torch
        out: list = [None] * len(self.parts)
        err: list = []

        def work(k):
            try:
                with torch.cuda.stream(self.streams[k]):
                    out[k] = fn(k, self.runners[k])
                    self.streams[k].synchronize()
            except BaseException as e:  # re-raised in the caller's thread
                err.append(e)

        if len(self.parts) == 1:
            work(0)
        else:
            threads = [threading.Thread(target=work, args=(k,)) for k in range(len(self.parts))]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
        if err:
            raise err[0]
        return out


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
                      predictor: str = "euler", groups: int = 1) -> BatchResult:
    """Advance every film from t_start to t_stop (mixed precision, lockstep).

    The Newton/TGCR parameters mean what they mean in
    :func:`nanowire_tdgl_amd.solvers.integrators.trapezoidal`; ``tune`` runs
    the stencil autotuner first; ``max_attempts`` bounds the number of
    lockstep attempts (None: until every film reaches t_stop); ``precond`` as
    in ``trapezoidal`` (each film takes its own automatic degree);
    ``predictor`` as in ``trapezoidal``, per film.

    ``groups`` > 1 runs the films as that many concurrent sub-batches
    (:class:`BatchGroups`, balanced by nodes; ``runner`` may instead be a
    BatchGroups built for these films, such as one grouped by ``keys`` measured
    on an earlier run): every film's result is the same bit for bit, the lockstep and
    ``max_attempts`` hold per group, ``attempts`` is the largest group's count
    and ``log[a][f]`` is None for a film whose group made fewer attempts."""
    if predictor not in ("euler", "extrapolate"):
        raise ValueError("predictor must be 'euler' or 'extrapolate'")
    opts = _native.solver_opts(newton_tol_f, newton_tol_dx, newton_max_iter, tol_gcr, 1e-4,
                               tgcr_max_iter, precision="mixed", mixed_tol_floor=mixed_tol_floor,
                               precond=precond)
    kw = dict(t_start=t_start, t_stop=t_stop, dt=dt, dt_min=dt_min, max_attempts=max_attempts,
              keep_log=keep_log, predictor=predictor, opts=opts)
    if groups == 1 and not isinstance(runner, BatchGroups):
        runner = runner or BatchRunner([f.device for f in films])
        X = runner.pack([f.x0 for f in films])
        if tune:
            runner.tune()
        r = _advance(runner, X, list(range(len(films))), **kw)
        r.states = [runner.unpack(X, f) for f in range(len(films))]
        return r
    if runner is not None and not isinstance(runner, BatchGroups):
        raise ValueError("with groups > 1, runner must be a BatchGroups")
    grp = runner or BatchGroups([f.device for f in films], groups)
    if sorted(i for part in grp.parts for i in part) != list(range(len(films))):
        raise ValueError(f"the BatchGroups runner holds {sum(map(len, grp.parts))} films; "
                         f"{len(films)} were given")
    import torch
    Xs = []
    for part, r, st in zip(grp.parts, grp.runners, grp.streams):
        with torch.cuda.stream(st):
            Xs.append(r.pack([films[i].x0 for i in part]))
    torch.cuda.synchronize()
    if tune:  # the groups timed together
        grp.tune()
    res = grp.run(lambda k, r: _advance(r, Xs[k], grp.parts[k], **kw))
    F = len(films)
    states, times, steps = [None] * F, [0.0] * F, [0] * F
    attempts = max(r.attempts for r in res)
    log = [[None] * F for _ in range(attempts)] if keep_log else []
    for part, r, run, X in zip(grp.parts, res, grp.runners, Xs):
        for j, f in enumerate(part):
            states[f], times[f], steps[f] = run.unpack(X, j), r.times[j], r.steps[j]
            for a, row in enumerate(r.log):
                log[a][f] = row[j]
    return BatchResult(states, times, steps, attempts, log)


def _advance(runner: BatchRunner, X, ids, *, t_start, t_stop, dt, dt_min, max_attempts, keep_log,
             predictor, opts) -> BatchResult:
    """The lockstep attempts of trapezoidal_batch over one tdgl_batch, on its
    packed state X (advanced in place); ``ids`` are the films' numbers in the
    caller's list (for messages).  The result's states are left empty."""
    F = runner.batch.nfilm
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
                    raise RuntimeError(f"film {ids[f]}: dt reduced below dt_min={dt_min:.1e} "
                                       f"at t={t[f]:.6e}")
    return BatchResult([], list(t), list(steps), attempts, log)
