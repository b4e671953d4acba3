"""Synthetic workloads of the benchmark (random draws from numpy's
default_rng on the host, in the reference's interior order; complex noise is
standard_normal(n) + 1j*standard_normal(n), real parts drawn first).

* ``film128_hole``: 128x128x16 cells, h = 0.5, kappa = 4.0, Bz = 0.4, a centred
  square hole (24..40)^2 through all z; psi = 1 + 1e-3 noise (seed 1), phi = 0.
* ``film256_spheres``: 256x256x64 cells, h = 0.5, kappa = 3.0, Bz = 0.5,
  insulating spheres of radius U[2, 4] cells with uniform centres (seed 2)
  until >= 1% of the interior nodes are insulating; psi = (1 + 1e-3 noise)
  (seed 2) times the superconducting mask, phi = 0.
* ``batch_films``: films from seed 3: Nx, Ny uniform integers in [32, 160],
  Nz in [4, 24], h = 0.5, kappa ~ U[1, 5], Bz ~ U[0.05, 0.8]; even-indexed
  films get one centred rectangular hole with sides U[0.1, 0.3] * Nx cells
  (in cell units: sides * h); psi = 1 + 1e-3 noise (seed 3 + 1000 + f), phi = 0.
"""

from __future__ import annotations

import numpy as np

from .core.device import Device
from .core.material import random_spheres
from .core.parameters import SimulationParameters
from .physics.applied_field import AppliedField
from .solvers.batch import BatchFilm


def _noise_state(n: int, seed: int, amp: float, mask=None) -> np.ndarray:
    rng = np.random.default_rng(seed)
    psi = 1.0 + amp * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    if mask is not None:
        psi = psi * mask
    return np.concatenate([psi, np.zeros(3 * n, complex)])


def film128_hole():
    p = SimulationParameters(Nx=128, Ny=128, Nz=16, hx=0.5, hy=0.5, hz=0.5, kappa=4.0)
    dev = Device(p, applied_field=AppliedField(Bz=0.4))
    dev.add_hole([(24.0, 24.0), (40.0, 24.0), (40.0, 40.0), (24.0, 40.0)])
    return dev, _noise_state(p.n_interior, 1, 1e-3)


def film256_spheres(Nx: int = 256, Ny: int = 256, Nz: int = 64, fraction: float = 0.01):
    p = SimulationParameters(Nx=Nx, Ny=Ny, Nz=Nz, hx=0.5, hy=0.5, hz=0.5, kappa=3.0)
    dev = Device(p, applied_field=AppliedField(Bz=0.5))
    centers, radii = random_spheres(p, fraction, 2.0, 4.0, seed=2)
    dev.add_spheres(centers, radii)
    mask = dev.material.interior_sc_mask
    return dev, _noise_state(p.n_interior, 2, 1e-3, mask)


def batch_film_specs(nfilm: int = 64, seed: int = 3, nmin: int = 32, nmax: int = 160,
                     zmin: int = 4, zmax: int = 24) -> list[dict]:
    """The batch's films as plain data: grid parameters, Bz, the hole polygon
    (or None) and the seed of the initial state's noise."""
    rng = np.random.default_rng(seed)
    specs = []
    for f in range(nfilm):
        Nx, Ny = (int(v) for v in rng.integers(nmin, nmax + 1, size=2))
        Nz = int(rng.integers(zmin, zmax + 1))
        kappa = float(rng.uniform(1.0, 5.0))
        bz = float(rng.uniform(0.05, 0.8))
        hole = None
        if f % 2 == 0:
            sx, sy = rng.uniform(0.1, 0.3, size=2) * Nx * 0.5
            cx, cy = Nx * 0.5 / 2.0, Ny * 0.5 / 2.0
            hole = [(cx - sx / 2, cy - sy / 2), (cx + sx / 2, cy - sy / 2),
                    (cx + sx / 2, cy + sy / 2), (cx - sx / 2, cy + sy / 2)]
        specs.append({"params": dict(Nx=Nx, Ny=Ny, Nz=Nz, hx=0.5, hy=0.5, hz=0.5, kappa=kappa),
                      "bz": bz, "hole": hole, "state_seed": seed + 1000 + f})
    return specs


def batch_state(spec: dict) -> np.ndarray:
    """Initial state of one batch film: psi = 1 + 1e-3 noise, phi = 0."""
    p = spec["params"]
    n = (p["Nx"] - 1) * (p["Ny"] - 1) * (p["Nz"] - 1)
    return _noise_state(n, spec["state_seed"], 1e-3)


def batch_films(nfilm: int = 64, seed: int = 3, nmin: int = 32, nmax: int = 160,
                zmin: int = 4, zmax: int = 24) -> list[BatchFilm]:
    films = []
    for spec in batch_film_specs(nfilm, seed, nmin, nmax, zmin, zmax):
        p = SimulationParameters(**spec["params"])
        dev = Device(p, applied_field=AppliedField(Bz=spec["bz"]))
        if spec["hole"]:
            dev.add_hole(spec["hole"])
        films.append(BatchFilm(dev, batch_state(spec)))
    return films
