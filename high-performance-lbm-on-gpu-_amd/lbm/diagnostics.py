"""Macroscopic snapshots and derived quantities (the interface of the reference's
``src/lbm/diagnostics.py:25-105``).

Solvers store post-collision populations f*; their moments give the state at
the time level of the collision that produced them::

    rho = 1 + sum_i f~*_i,    rho u = sum_i c_i f*_i - F/2

The HIP backend computes rho and u on the device (``lbm_macroscopic``); a
``MacroscopicFields`` holds the host copy. The velocity is kept as one read-only
(d, *shape) array; ``u`` hands out its components.
"""

from __future__ import annotations

import numpy as np

_AXES = "xyz"


class MacroscopicFields:
    """Read-only host copies of rho and u (solid nodes report rho = 1, u = 0).

    Built as ``MacroscopicFields(rho=..., u=(ux, uy[, uz]), solid=...)``; exposes
    ``rho``, ``u`` (tuple of components), ``solid``, ``ux``/``uy``/``uz``, ``d``,
    ``speed`` and ``total_mass``.
    """

    __slots__ = ("rho", "solid", "u", "_velocity")

    def __init__(self, rho: np.ndarray, u, solid: np.ndarray):
        velocity = np.stack([np.asarray(c) for c in u])
        for array in (rho, solid, velocity):
            array.setflags(write=False)
        self.rho, self.solid, self._velocity = rho, solid, velocity
        self.u = tuple(velocity)  # one read-only view per component, made once

    @property
    def d(self) -> int:
        return self._velocity.shape[0]

    def __getattr__(self, name: str):
        # ux, uy, uz: velocity components by axis letter
        if len(name) == 2 and name[0] == "u" and name[1] in _AXES:
            axis = _AXES.index(name[1])
            if axis >= self.d:
                raise AttributeError(f"{self.d}D fields have no {name} component")
            return self.u[axis]
        raise AttributeError(name)

    @property
    def speed(self) -> np.ndarray:
        return np.sqrt(np.square(self._velocity).sum(axis=0))

    @property
    def total_mass(self) -> float:
        return float(np.sum(self.rho, where=~self.solid, dtype=np.float64))


def vorticity(*u: np.ndarray) -> np.ndarray:
    """Central-difference curl of the velocity, arrays indexed [y, x] / [z, y, x]: the
    scalar d(uy)/dx - d(ux)/dy in 2D, the (3, ...) vector in 3D."""
    d = len(u)
    if d not in (2, 3):
        raise ValueError(f"vorticity needs 2 or 3 velocity components; got {d}")

    def derivative(component: int, along: int) -> np.ndarray:
        # array axis of spatial direction `along` (x is the last array axis)
        return np.gradient(u[component], axis=d - 1 - along)

    def curl(a: int) -> np.ndarray:
        b, c = (a + 1) % 3, (a + 2) % 3
        return derivative(c, b) - derivative(b, c)

    return curl(2) if d == 2 else np.stack([curl(a) for a in range(3)])


def kinetic_energy(rho: np.ndarray, *u: np.ndarray) -> float:
    """Total kinetic energy, half the float64 sum of rho |u|^2."""
    u_squared = np.add.reduce([np.square(c) for c in u])
    return float(0.5 * np.sum(rho * u_squared, dtype=np.float64))
