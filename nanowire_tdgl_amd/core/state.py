"""Flat state vector [psi, phi_x, phi_y, (phi_z)] with named views
(mirror of tdgl3d's StateVector, src/tdgl3d/core/state.py:11-121)."""

from __future__ import annotations

import numpy as np

from .parameters import SimulationParameters


class StateVector:
    def __init__(self, data, params: SimulationParameters):
        expected = params.n_state
        if data.shape != (expected,):
            raise ValueError(
                f"State vector length {data.shape} does not match expected ({expected},) "
                f"for grid {params.Nx}×{params.Ny}×{params.Nz}.")
        self._data = np.asarray(data, dtype=np.complex128)
        self._params = params

    def _block(self, b: int) -> slice:
        n = self._params.n_interior
        return slice(b * n, (b + 1) * n)

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def psi(self) -> np.ndarray:
        return self._data[self._block(0)]

    @psi.setter
    def psi(self, val) -> None:
        self._data[self._block(0)] = val

    @property
    def phi_x(self) -> np.ndarray:
        return self._data[self._block(1)]

    @phi_x.setter
    def phi_x(self, val) -> None:
        self._data[self._block(1)] = val

    @property
    def phi_y(self) -> np.ndarray:
        return self._data[self._block(2)]

    @phi_y.setter
    def phi_y(self, val) -> None:
        self._data[self._block(2)] = val

    @property
    def phi_z(self) -> np.ndarray:
        if not self._params.is_3d:
            raise AttributeError("phi_z is not available for 2-D simulations (Nz == 1).")
        return self._data[self._block(3)]

    @phi_z.setter
    def phi_z(self, val) -> None:
        if not self._params.is_3d:
            raise AttributeError("phi_z is not available for 2-D simulations.")
        self._data[self._block(3)] = val

    def copy(self) -> "StateVector":
        return StateVector(self._data.copy(), self._params)

    @classmethod
    def uniform_superconducting(cls, params: SimulationParameters) -> "StateVector":
        data = np.zeros(params.n_state, dtype=np.complex128)
        data[: params.n_interior] = 1.0
        return cls(data, params)

    @classmethod
    def from_components(cls, psi, phi_x, phi_y, phi_z, params: SimulationParameters) -> "StateVector":
        parts = [np.ravel(psi), np.ravel(phi_x), np.ravel(phi_y)]
        if params.is_3d:
            if phi_z is None:
                raise ValueError("phi_z required for 3-D.")
            parts.append(np.ravel(phi_z))
        return cls(np.concatenate(parts).astype(np.complex128), params)

    def __repr__(self) -> str:
        p = self._params
        return (f"StateVector(Nx={p.Nx}, Ny={p.Ny}, Nz={p.Nz}, n_interior={p.n_interior}, "
                f"|ψ|_max={np.max(np.abs(self.psi)):.4f})")
