"""Macroscopic snapshots and derived quantities (``src/lbm/diagnostics.py:25-105`` of the reference).

Solvers store post-collision populations f*; their moments give the state at
the time level of the collision that produced them::

    rho = 1 + sum_i f~*_i,    rho u = sum_i c_i f*_i - F/2

The HIP backend computes rho and u on the device (``lbm_macroscopic``);
these host classes wrap the result.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True, eq=False)
class MacroscopicFields:
    """Read-only host copies of rho and u (solid nodes report rho = 1, u = 0)."""

    rho: np.ndarray
    u: tuple[np.ndarray, ...]
    solid: np.ndarray

    def __post_init__(self) -> None:
        object.__setattr__(self, "u", tuple(self.u))
        for array in (self.rho, self.solid, *self.u):
            array.setflags(write=False)

    @property
    def d(self) -> int:
        return len(self.u)

    @property
    def ux(self) -> np.ndarray:
        return self.u[0]

    @property
    def uy(self) -> np.ndarray:
        return self.u[1]

    @property
    def uz(self) -> np.ndarray:
        if self.d < 3:
            raise AttributeError("2D fields have no uz component")
        return self.u[2]

    @property
    def speed(self) -> np.ndarray:
        return np.sqrt(sum(c * c for c in self.u))

    @property
    def total_mass(self) -> float:
        return float(self.rho[~self.solid].sum(dtype=np.float64))


def vorticity(*u: np.ndarray) -> np.ndarray:
    """Central-difference vorticity: scalar in 2D, (3, ...) vector in 3D."""
    if len(u) == 2:
        ux, uy = u
        return np.gradient(uy, axis=1) - np.gradient(ux, axis=0)
    if len(u) == 3:
        ux, uy, uz = u
        dz, dy, dx = (lambda f, a=a: np.gradient(f, axis=a) for a in range(3))
        return np.stack([dy(uz) - dz(uy), dz(ux) - dx(uz), dx(uy) - dy(ux)])
    raise ValueError(f"vorticity needs 2 or 3 velocity components; got {len(u)}")


def kinetic_energy(rho: np.ndarray, *u: np.ndarray) -> float:
    return float(0.5 * np.sum(rho * sum(c * c for c in u), dtype=np.float64))
