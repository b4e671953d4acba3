"""Builders that turn the shared case specs into oracle films or library devices."""

from __future__ import annotations

import numpy as np

from oracle import tdgl_oracle as O


def oracle_film(case) -> O.Film:
    p = case["p"]
    g = O.Grid(p["Nx"], p["Ny"], p["Nz"], p.get("hx", 1.0), p.get("hy", 1.0), p.get("hz", 1.0),
               p.get("kappa", 5.0), p.get("periodic_x", False), p.get("periodic_y", False),
               p.get("periodic_z", False))
    f = O.Film(g, case.get("trilayer"))
    for verts, zr in case["holes"]:
        f.add_hole(verts, zr)
    return f


def device(case, **kw):
    import nanowire_tdgl_amd as T
    p = T.SimulationParameters(**case["p"])
    tri = None
    if case.get("trilayer"):
        b, i, t = case["trilayer"]
        tri = T.Trilayer(T.Layer(*b), T.Layer(*i), T.Layer(*t))
    dev = T.Device(p, trilayer=tri, **kw)
    for verts, zr in case["holes"]:
        dev.add_hole(verts, zr)
    return dev


def rand_state(n, nphi, seed, psi_noise=0.3, phi_scale=0.2):
    rng = np.random.default_rng(seed)
    psi = 1.0 + psi_noise * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    phi = phi_scale * rng.standard_normal(nphi * n)
    return np.concatenate([psi, phi.astype(np.complex128)])


def rel_err(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale
