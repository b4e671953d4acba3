"""Solution container (mirror of tdgl3d's core/solution.py:19-549).

The check-path observables B = curl(phi) (with the boundary conditions at the
saved time) and J_s (on the expanded state without them) are computed by the
device kernel ``tdgl_observables``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from ..mesh.indices import GridIndices
from .parameters import SimulationParameters


@dataclass
class Solution:
    times: np.ndarray
    states: np.ndarray
    params: SimulationParameters
    idx: GridIndices
    device: Optional[object] = None
    metadata: Optional[dict] = None

    @property
    def n_steps(self) -> int:
        return self.states.shape[1]

    def _block(self, b: int, step: int) -> np.ndarray:
        n = self.params.n_interior
        return self.states[b * n:(b + 1) * n, step]

    def psi(self, step: int = -1) -> np.ndarray:
        return self._block(0, step)

    def phase(self, step: int = -1, mask_threshold: float = 0.02) -> np.ndarray:
        psi = self.psi(step)
        ph = np.angle(psi).astype(np.float64)
        ph[np.abs(psi) ** 2 < mask_threshold] = np.nan
        return ph

    def psi_squared(self, step: int = -1) -> np.ndarray:
        return np.abs(self.psi(step)) ** 2

    def phi_x(self, step: int = -1) -> np.ndarray:
        return self._block(1, step)

    def phi_y(self, step: int = -1) -> np.ndarray:
        return self._block(2, step)

    def phi_z(self, step: int = -1) -> np.ndarray:
        if not self.params.is_3d:
            raise AttributeError("No phi_z for 2-D simulations.")
        return self._block(3, step)

    def _material(self):
        return getattr(self.device, "material", None) if self.device is not None else None

    def bfield(self, step: int = -1, full_interior: bool = True):
        """B = curl(phi) at a saved step (solution.py:94-158).  full_interior:
        every interior node, with the BCs of the applied field at times[step]
        (t_stop = times[-1]); otherwise the one-layer-in subset of eval_bfield."""
        p = self.params
        state = self.states[:, step]
        if not full_interior:
            return _bfield_subset(state, p, self.idx)
        from ..physics.applied_field import build_boundary_field_vectors
        from ..physics.rhs import BoundaryVectors, device_observables
        if self.device is None or getattr(self.device, "applied_field", None) is None:
            raise NotImplementedError("bfield(full_interior=True) needs the Solution's device")
        bx, by, bz = self.device.applied_field.evaluate(self.times[step], self.times[-1])
        u = BoundaryVectors.from_field(build_boundary_field_vectors(bx, by, bz, p, self.idx))
        b, _ = device_observables(state, p, self.idx, u, self._material())
        return b[0], b[1], b[2]

    def _reshape_interior(self, arr: np.ndarray, slice_z: int = 0) -> np.ndarray:
        p = self.params
        if p.is_3d:
            return arr.reshape(p.Nx - 1, p.Ny - 1, max(p.Nz - 1, 1))[:, :, slice_z]
        return arr.reshape(p.Nx - 1, p.Ny - 1)

    def psi_squared_2d(self, step: int = -1, slice_z: int = 0) -> np.ndarray:
        return np.abs(self._reshape_interior(self.psi(step), slice_z)) ** 2

    def supercurrent_density(self, step: int = -1):
        """J_s = Im[conj(psi) e^{-i phi} psi(+a)] on the expanded state without
        BCs (solution.py:180-217)."""
        from ..physics.rhs import BoundaryVectors, device_observables
        p = self.params
        z = np.zeros(p.dim_x)
        _, j = device_observables(self.states[:, step], p, self.idx,
                                  BoundaryVectors(z, z, z, uniform=(0.0, 0.0, 0.0)),
                                  self._material())
        return j[0], j[1], j[2]

    def save(self, filename: str) -> None:
        """HDF5 layout of solution.py:400-461 (needs h5py)."""
        import h5py
        with h5py.File(filename, "w") as f:
            f.create_dataset("times", data=self.times)
            f.create_dataset("states", data=self.states)
            g = f.create_group("params")
            for key in ("Nx", "Ny", "Nz", "hx", "hy", "hz", "kappa"):
                g.attrs[key] = getattr(self.params, key)
            gi = f.create_group("idx")
            for key in self.idx.array_names():
                gi.create_dataset(key, data=getattr(self.idx, key))

    @staticmethod
    def load(filename: str) -> "Solution":
        import h5py
        with h5py.File(filename, "r") as f:
            times, states = f["times"][:], f["states"][:]
            a = f["params"].attrs
            params = SimulationParameters(Nx=int(a["Nx"]), Ny=int(a["Ny"]), Nz=int(a["Nz"]),
                                          hx=float(a["hx"]), hy=float(a["hy"]), hz=float(a["hz"]),
                                          kappa=float(a["kappa"]))
            idx = GridIndices(**{k: f["idx"][k][:] for k in f["idx"].keys()})
        return Solution(times=times, states=states, params=params, idx=idx)


def _bfield_subset(state, p: SimulationParameters, idx: GridIndices):
    """eval_bfield(full_interior=False) (bfield.py:81-151): curl on the subset
    `bfield_interior`, which indexes the interior vector as if it were i-fastest."""
    n = p.n_interior
    mj, mk = p.Nx - 1, (p.Nx - 1) * (p.Ny - 1)
    px, py = state[n:2 * n], state[2 * n:3 * n]
    m = idx.bfield_interior
    if p.is_3d:
        pz = state[3 * n:4 * n]
        bx = (1.0 / (p.hy * p.hz)) * (py[m] - py[m + mk] - pz[m] + pz[m + mj])
        by = (1.0 / (p.hz * p.hx)) * (pz[m] - pz[m + 1] - px[m] + px[m + mk])
    else:
        bx = np.zeros(len(m))
        by = np.zeros(len(m))
    bz = (1.0 / (p.hx * p.hy)) * (px[m] - px[m + mj] - py[m] + py[m + 1])
    return np.real(bx), np.real(by), np.real(bz)
