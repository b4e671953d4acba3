"""CPU baseline sampler (TEST/BENCH INFRASTRUCTURE ONLY; run by bench.py's
cpu_baseline leg as a subprocess).

Times the oracle's faithful FP64 restatement of tdgl3d's trapezoidal step
(CSR operators rebuilt per eval_f, TGCR with FD-JVP and MGS, Newton) on the
first step of the 128x128x16 film with the 32x32 hole, for a bounded wall
time, and prints one JSON line with the elapsed time and the number of
eval_f calls completed.  bench.py scales that to a full step with the step's
measured Newton/TGCR counts (eval_f calls per step = 1 + sum_newton (2 + 2 k)).

    OMP_NUM_THREADS=1 python -m oracle.cpu_sample --budget 20
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle import tdgl_oracle as O  # noqa: E402


class _Budget(Exception):
    pass


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--budget", type=float, default=20.0)
    ap.add_argument("--Nx", type=int, default=128)
    ap.add_argument("--Ny", type=int, default=128)
    ap.add_argument("--Nz", type=int, default=16)
    args = ap.parse_args()
    g = O.Grid(args.Nx, args.Ny, args.Nz, 0.5, 0.5, 0.5, 4.0)
    t0 = time.perf_counter()
    film = O.Film(g)
    c = 0.5 * args.Nx * 0.5  # hole centred, 32 cells wide at Nx = 128
    w = 8.0 * args.Nx / 128.0
    film.add_hole([(c - w, c - w), (c + w, c - w), (c + w, c + w), (c - w, c + w)])
    setup = time.perf_counter() - t0
    n = g.n_interior
    rng = np.random.default_rng(1)
    X = np.concatenate([1.0 + 1e-3 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)),
                        np.zeros(3 * n, complex)])
    u = O.boundary_vectors(0.0, 0.0, 0.0, g, film.idx)  # the field is off at t = 0
    calls = [0]
    start = time.perf_counter()

    def f(Y):
        if time.perf_counter() - start > args.budget:
            raise _Budget()
        calls[0] += 1
        return O.eval_f_csr(Y, g, film.idx, u, film.material)

    finished = False
    try:
        fn = f(X)
        dt = 0.2
        O.newton_trap(f, X + dt * fn, X + (dt / 2.0) * fn, dt, tol_gcr=1e-8)
        finished = True
    except _Budget:
        pass
    el = time.perf_counter() - start
    print(json.dumps({"elapsed_s": el, "evals": calls[0], "finished_step": finished,
                      "setup_s": setup, "n_interior": n,
                      "threads": os.environ.get("OMP_NUM_THREADS", "unset")}))


if __name__ == "__main__":
    main()
