"""Supercurrent density on full-grid arrays (mirror of tdgl3d's
physics/current_density.py:14-72): J_s,a = Im[conj(psi) e^{-i phi_a} psi(+a)]
at the interior nodes of the expanded state (no boundary conditions).  The
device computes the same quantity for Solution.supercurrent_density."""

from __future__ import annotations

import numpy as np


def eval_supercurrent_density(psi_full, phi_x_full, phi_y_full, phi_z_full, params, idx):
    m = idx.interior_to_full
    jx = np.imag(np.conj(psi_full[m]) * np.exp(-1j * phi_x_full[m]) * psi_full[m + 1])
    jy = np.imag(np.conj(psi_full[m]) * np.exp(-1j * phi_y_full[m]) * psi_full[m + params.mj])
    if params.is_3d:
        jz = np.imag(np.conj(psi_full[m]) * np.exp(-1j * phi_z_full[m]) * psi_full[m + params.mk])
    else:
        jz = np.zeros(params.n_interior)
    return jx, jy, jz
