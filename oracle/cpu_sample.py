"""CPU baseline sampler (TEST/BENCH INFRASTRUCTURE ONLY; run by bench.py's
cpu_baseline leg as a subprocess).

Times the oracle's faithful FP64 restatement of tdgl3d's trapezoidal step
(CSR operators rebuilt per eval_f, TGCR with FD-JVP and MGS, Newton, the
dt/10 retry; tdgl3d's integrators.py:150-197) on whole steps of the films of
the 64-film batch (seed 3), smallest films first, for a bounded wall time:

* ``--workers 1``: one process, films one after the other (one thread: the
  caller sets OMP_NUM_THREADS=1; scipy's sparse products are single-threaded);
* ``--workers W``: W processes, one film step each at a time (the way a CPU
  user runs independent films on all cores).

Each film takes one trapezoidal step of dt = 0.2 from its initial state with
its applied field on (the batch's timed steps have the field on), TGCR
tolerance 1e-4 and Newton's defaults (the batch's settings).  A film started
before the budget ran out is finished, so the sample is whole steps only.
Prints one JSON line: films done, interior nodes advanced, wall seconds,
cell-steps/s.

    OMP_NUM_THREADS=1 python -m oracle.cpu_sample --budget 20 --workers 1
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import FIRST_COMPLETED, ProcessPoolExecutor, wait

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

DT = 0.2
TOL_GCR = 1e-4


def film_step(spec: dict) -> dict:
    """One accepted trapezoidal step of one batch film on the oracle."""
    from nanowire_tdgl_amd.workloads import batch_state
    from oracle import tdgl_oracle as O
    p = spec["params"]
    g = O.Grid(p["Nx"], p["Ny"], p["Nz"], p["hx"], p["hy"], p["hz"], p["kappa"])
    film = O.Film(g)
    if spec["hole"]:
        film.add_hole(spec["hole"])
    X = batch_state(spec)
    u = O.boundary_vectors(0.0, 0.0, spec["bz"], g, film.idx)
    calls = [0]

    def f(Y):
        calls[0] += 1
        return O.eval_f_csr(Y, g, film.idx, u, film.material)

    t0 = time.perf_counter()
    dt, newton = DT, 0
    while True:  # integrators.py:151-197 for one step
        fn = f(X)
        _, conv, it = O.newton_trap(f, X + dt * fn, X + (dt / 2.0) * fn, dt, tol_gcr=TOL_GCR)
        newton += it
        if conv:
            break
        dt /= 10.0
        if dt < 1e-8:
            raise RuntimeError("oracle step did not converge")
    return {"n": g.n_interior, "s": time.perf_counter() - t0, "evals": calls[0], "newton": newton}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--budget", type=float, default=20.0)
    ap.add_argument("--workers", type=int, default=1)
    args = ap.parse_args()
    from nanowire_tdgl_amd.workloads import batch_film_specs
    specs = sorted(batch_film_specs(), key=lambda s: (s["params"]["Nx"] - 1) * (s["params"]["Ny"] - 1)
                   * (s["params"]["Nz"] - 1))
    done = []
    start = time.perf_counter()
    if args.workers <= 1:
        for spec in specs:
            if time.perf_counter() - start > args.budget:
                break
            done.append(film_step(spec))
    else:
        with ProcessPoolExecutor(max_workers=args.workers) as pool:
            pending = set()
            it = iter(specs)
            for _ in range(args.workers):
                pending.add(pool.submit(film_step, next(it)))
            while pending:
                fin, pending = wait(pending, return_when=FIRST_COMPLETED)
                done += [x.result() for x in fin]
                for _ in fin:
                    if time.perf_counter() - start <= args.budget:
                        spec = next(it, None)
                        if spec is not None:
                            pending.add(pool.submit(film_step, spec))
    el = time.perf_counter() - start
    nodes = sum(d["n"] for d in done)
    print(json.dumps({"workers": max(args.workers, 1), "films": len(done), "cell_steps": nodes,
                      "elapsed_s": el, "value": nodes / el,
                      "film_sizes": [d["n"] for d in done],
                      "film_step_s": [round(d["s"], 3) for d in done],
                      "newton_iters": [d["newton"] for d in done],
                      "threads_per_worker": os.environ.get("OMP_NUM_THREADS", "unset")}))


if __name__ == "__main__":
    main()
