"""Steady-state detection on saved states (mirror of tdgl3d's
analysis/convergence.py:14-197).

Kept as the reference has it: the supercurrent part calls
eval_supercurrent_density(solution, device, step=...), which does not match
that function's signature (convergence.py:88, physics/current_density.py:14-21);
the TypeError is caught and recorded under 'current_error', so the |J_s|
metrics never appear."""

from __future__ import annotations

import numpy as np


def compute_convergence_metrics(solution, device=None, step: int = -1, window_size: int = 10) -> dict:
    metrics = {}
    if step < 0:
        step = solution.n_steps + step
    past = max(0, step - window_size)
    p2c = np.abs(solution.psi(step=step)) ** 2
    p2p = np.abs(solution.psi(step=past)) ** 2
    if device is not None and device.material is not None:
        sc = device.material.interior_sc_mask
        mc, mp = float(np.mean(p2c[sc > 0])), float(np.mean(p2p[sc > 0]))
    else:
        mc, mp = float(np.mean(p2c)), float(np.mean(p2p))
    rel = abs(mc - mp) / mc if mc > 1e-10 else float("inf")
    metrics["psi2_mean_current"] = mc
    metrics["psi2_mean_past"] = mp
    metrics["psi2_rel_change"] = rel
    if device is not None:
        try:
            from ..physics.current_density import eval_supercurrent_density
            jx, jy, jz = eval_supercurrent_density(solution, device, step=step)
            jmag = np.sqrt(jx ** 2 + jy ** 2 + (jz ** 2 if jz is not None else 0))
            jx, jy, jz = eval_supercurrent_density(solution, device, step=past)
            jmag_p = np.sqrt(jx ** 2 + jy ** 2 + (jz ** 2 if jz is not None else 0))
            if device.material is not None:
                sc = device.material.interior_sc_mask
                cc, cp = float(np.mean(jmag[sc > 0])), float(np.mean(jmag_p[sc > 0]))
            else:
                cc, cp = float(np.mean(jmag)), float(np.mean(jmag_p))
            metrics["current_mean_current"] = cc
            metrics["current_mean_past"] = cp
            metrics["current_rel_change"] = abs(cc - cp) / cc if cc > 1e-10 else float("inf")
        except Exception as e:  # convergence.py:113-115
            metrics["current_error"] = str(e)
    return metrics


def check_steady_state(solution, device=None, window_size: int = 10, psi_threshold: float = 1e-4,
                       current_threshold: float = 1e-4, start_step: int = 20):
    n_steps = solution.n_steps
    if n_steps < start_step + window_size:
        return False, -1, compute_convergence_metrics(solution, device, step=-1,
                                                      window_size=window_size)
    steady = -1
    for step in range(start_step, n_steps):
        m = compute_convergence_metrics(solution, device, step=step, window_size=window_size)
        ok = m["psi2_rel_change"] < psi_threshold
        if device is not None and "current_rel_change" in m:
            ok = ok and m["current_rel_change"] < current_threshold
        if ok:
            steady = step
            break
    final = compute_convergence_metrics(solution, device, step=-1, window_size=window_size)
    if steady >= 0:
        final["steady_time"] = float(solution.times[steady])
        final["steady_step"] = steady
        return True, steady, final
    return False, -1, final
