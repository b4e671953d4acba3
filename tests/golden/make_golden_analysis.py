"""Write the golden fixtures of the post-processing functions (tdgl3d's
analysis package: vortex_counting.py, convergence.py).

Run in the development container only (the GPU box has no copy of tdgl3d):

    python -B tests/golden/make_golden_analysis.py /root/reference/src

It imports the reference package from the given source directory, applies its
analysis functions to small synthetic saved trajectories built here (vortex
lattices with noise, relaxing amplitudes, a film with a hole) and writes
``analysis.npz`` (inputs and outputs only) next to this script.
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def vortex_states(Nx, Ny, Nz, n_saved, seed):
    """psi = prod_k tanh(r_k / 2) exp(i s_k theta_k) (+ noise), vortex centres
    drifting with the saved step; phi random (real)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = Nx - 1, Ny - 1, max(Nz - 1, 1)
    n = nx * ny * nz
    nphi = 3 if Nz > 1 else 2
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    centres = [(0.3 * nx, 0.35 * ny, 1), (0.7 * nx, 0.6 * ny, -1), (0.45 * nx, 0.8 * ny, 1)]
    states = np.zeros(((1 + nphi) * n, n_saved), complex)
    for s in range(n_saved):
        psi = np.ones((nx, ny, nz), complex)
        for cx, cy, sgn in centres:
            dx, dy = I - (cx + 0.13 * s), J - (cy - 0.07 * s) + 0.1 * K
            psi = psi * np.tanh(np.hypot(dx, dy) / 2.0) * np.exp(1j * sgn * np.arctan2(dy, dx))
        psi = psi * (1.0 - 0.4 * np.exp(-s / 4.0))
        psi = psi + 0.02 * (rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape))
        states[:n, s] = psi.ravel()  # reference interior order: i slowest, k fastest
        states[n:, s] = 0.1 * rng.standard_normal(nphi * n)
    return states


CASES = {"v3d": (24, 20, 5, 0.3, None), "v2d": (22, 18, 1, 0.2, None),
         "vhole": (26, 22, 4, 0.4, [(6.0, 6.0), (12.0, 6.0), (12.0, 11.0), (6.0, 11.0)])}
N_SAVED = 32


def case_seed(name: str, Nx: int) -> int:
    return len(name) * 7 + Nx


def main(src: str) -> None:
    sys.path.insert(0, src)
    import tdgl3d
    from tdgl3d.analysis.convergence import check_steady_state, compute_convergence_metrics
    from tdgl3d.analysis.vortex_counting import (count_hole_flux_quanta, count_vortices_plaquette,
                                                 count_vortices_polygon, find_vortex_cores)
    from tdgl3d.core.solution import Solution

    out = {}
    for name, (Nx, Ny, Nz, bz, hole) in CASES.items():
        p = tdgl3d.SimulationParameters(Nx=Nx, Ny=Ny, Nz=Nz, hx=0.5, hy=0.5, hz=0.5, kappa=2.0)
        dev = tdgl3d.Device(p, applied_field=tdgl3d.AppliedField(Bz=bz, t_on_fraction=1.0))
        if hole:
            dev.add_hole(hole)
        n_saved = N_SAVED
        st = vortex_states(Nx, Ny, Nz, n_saved, seed=case_seed(name, Nx))
        times = np.linspace(0.0, 6.2, n_saved)
        sol = Solution(times=times, states=st, params=p, idx=dev.idx, device=dev)
        # the states are regenerated by the tests (vortex_states is deterministic);
        # their digest pins the inputs
        out[f"{name}_states_sha"] = np.array(hashlib.sha256(st.tobytes()).hexdigest())
        out[f"{name}_times"] = times
        nz_int = max(Nz - 1, 1)
        for sz in range(min(nz_int, 2)):
            for thr in (0.8, 0.5):
                nv, pos, wn = count_vortices_plaquette(sol, dev, slice_z=sz, step=-1,
                                                       winding_threshold=thr)
                out[f"{name}_plaq_{sz}_{thr}_n"] = np.array(nv)
                out[f"{name}_plaq_{sz}_{thr}_pos"] = pos
                out[f"{name}_plaq_{sz}_{thr}_wn"] = wn
            polys = [np.array([[3.0, 3.0], [15.0, 3.0], [15.0, 14.0], [3.0, 14.0]]),
                     np.array([[1.5, 2.5], [40.0, 2.0], [18.2, 17.9], [2.0, 16.0], [1.5, 2.5]]),
                     np.array([[-3.0, 5.0], [9.0, 5.0], [9.0, 30.0]])]
            for k, poly in enumerate(polys):
                out[f"{name}_poly_{sz}_{k}"] = np.array(count_vortices_polygon(sol, dev, poly, slice_z=sz, step=-1))
            out[f"{name}_cores_{sz}"] = find_vortex_cores(sol, dev, slice_z=sz, step=-1)
            out[f"{name}_cores2_{sz}"] = find_vortex_cores(sol, dev, slice_z=sz, step=5, threshold=0.3,
                                                           separation=3)
            out[f"{name}_flux_{sz}"] = np.array(count_hole_flux_quanta(sol, dev, (4, 14, 3.5, 12.7),
                                                                       slice_z=sz, step=-2))
        for w in (5, 10):
            m = compute_convergence_metrics(sol, dev, step=-3, window_size=w)
            out[f"{name}_conv_{w}_keys"] = np.array(sorted(m))
            out[f"{name}_conv_{w}_vals"] = np.array([str(m[k]) for k in sorted(m)])
            m = compute_convergence_metrics(sol, None, step=7, window_size=w)
            out[f"{name}_convnd_{w}_vals"] = np.array([m["psi2_mean_current"], m["psi2_mean_past"],
                                                      m["psi2_rel_change"]])
        for thr, start in ((1e-2, 5), (1e-4, 10), (3e-3, 2)):
            ok, s0, m = check_steady_state(sol, dev, window_size=4, psi_threshold=thr, start_step=start)
            out[f"{name}_steady_{thr}_{start}"] = np.array([int(ok), s0, m.get("steady_time", -1.0)])
    np.savez_compressed(HERE / "analysis.npz", **out)
    print(f"wrote {len(out)} arrays to analysis.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
