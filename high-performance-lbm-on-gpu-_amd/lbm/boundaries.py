"""Solid masks for simple wall geometries (``src/lbm/boundaries.py:107-142`` of the reference).

Halfway bounce-back itself lives in the stream-collide kernels: a population
pulled from a solid node x - c_i is replaced by f*_opp(i)(x) + 6 w_i (c_i . u_w),
with u_w the wall velocity stored at the solid node (Ladd 1994, rho_w = 1).
"""

from __future__ import annotations

import numpy as np


def channel_walls(nx: int, ny: int, nz: int | None = None) -> np.ndarray:
    """Solid first and last y rows; periodic in x (and z). Walls at y = 0.5 and ny - 1.5."""
    if nx <= 0 or ny < 3 or (nz is not None and nz <= 0):
        raise ValueError(f"channel needs nx > 0, ny >= 3, nz > 0; got nx={nx}, ny={ny}, nz={nz}")
    solid = np.zeros((ny, nx) if nz is None else (nz, ny, nx), dtype=bool)
    solid[..., 0, :] = True
    solid[..., -1, :] = True
    return solid


def cavity_walls(n: int, lid_velocity: float, dimensions: int = 2):
    """Square / cubic cavity of side n in a one-node solid frame; the top (y = n + 1) moves in +x."""
    if dimensions not in (2, 3):
        raise ValueError(f"cavity dimensions must be 2 or 3; got {dimensions}")
    if n <= 0:
        raise ValueError(f"cavity size must be positive; got n={n}")
    solid = np.ones((n + 2,) * dimensions, dtype=bool)
    solid[(slice(1, -1),) * dimensions] = False
    wall = np.zeros((dimensions, *solid.shape))
    wall[0, ..., -1, :] = lid_velocity
    return solid, wall
