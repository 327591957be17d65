"""Centreline velocities of the lid-driven cavity from Ghia, Ghia & Shin (1982).

U. Ghia, K. N. Ghia, C. T. Shin, "High-Re solutions for incompressible flow using the
Navier-Stokes equations and a multigrid method", J. Comput. Phys. 48 (1982) 387-411,
Tables I and II (129 x 129 grid for Re <= 1000). Positions are in units of the cavity
side, velocities in units of the lid speed. The same published table backs the
reference's ``src/lbm/cases/ghia1982.py:30-53``; ``ghia_centerlines`` keeps its API
(``src/lbm/cases/ghia1982.py:56-70``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Table I: u / U on the vertical centreline x = 0.5, from the lid (y = 1) down.
#            y        Re=100    Re=400    Re=1000
_TABLE_U = (
    (1.0000, 1.00000, 1.00000, 1.00000),
    (0.9766, 0.84123, 0.75837, 0.65928),
    (0.9688, 0.78871, 0.68439, 0.57492),
    (0.9609, 0.73722, 0.61756, 0.51117),
    (0.9531, 0.68717, 0.55892, 0.46604),
    (0.8516, 0.23151, 0.29093, 0.33304),
    (0.7344, 0.00332, 0.16256, 0.18719),
    (0.6172, -0.13641, 0.02135, 0.05702),
    (0.5000, -0.20581, -0.11477, -0.06080),
    (0.4531, -0.21090, -0.17119, -0.10648),
    (0.2813, -0.15662, -0.32726, -0.27805),
    (0.1719, -0.10150, -0.24299, -0.38289),
    (0.1016, -0.06434, -0.14612, -0.29730),
    (0.0703, -0.04775, -0.10338, -0.22220),
    (0.0625, -0.04192, -0.09266, -0.20196),
    (0.0547, -0.03717, -0.08186, -0.18109),
    (0.0000, 0.00000, 0.00000, 0.00000),
)  # fmt: skip

# Table II: v / U on the horizontal centreline y = 0.5, from the right wall (x = 1).
#            x        Re=100    Re=400    Re=1000
_TABLE_V = (
    (1.0000, 0.00000, 0.00000, 0.00000),
    (0.9688, -0.05906, -0.12146, -0.21388),
    (0.9609, -0.07391, -0.15663, -0.27669),
    (0.9531, -0.08864, -0.19254, -0.33714),
    (0.9453, -0.10313, -0.22847, -0.39188),
    (0.9063, -0.16914, -0.23827, -0.51550),
    (0.8594, -0.22445, -0.44993, -0.42665),
    (0.8047, -0.24533, -0.38598, -0.31966),
    (0.5000, 0.05454, 0.05186, 0.02526),
    (0.2344, 0.17527, 0.30174, 0.32235),
    (0.2266, 0.17507, 0.30203, 0.33075),
    (0.1563, 0.16077, 0.28124, 0.37095),
    (0.0938, 0.12317, 0.22965, 0.32627),
    (0.0781, 0.10890, 0.20920, 0.30353),
    (0.0703, 0.10091, 0.19713, 0.29012),
    (0.0625, 0.09233, 0.18360, 0.27485),
    (0.0000, 0.00000, 0.00000, 0.00000),
)  # fmt: skip

REYNOLDS_NUMBERS = (100, 400, 1000)


@dataclass(frozen=True)
class CenterlineData:
    y: np.ndarray
    """Positions along the vertical centreline x = 0.5."""
    u: np.ndarray
    """u / U at those positions."""
    x: np.ndarray
    """Positions along the horizontal centreline y = 0.5."""
    v: np.ndarray
    """v / U at those positions."""


def _read_only(values) -> np.ndarray:
    array = np.array(values, dtype=np.float64)
    array.setflags(write=False)
    return array


def ghia_centerlines(reynolds: int) -> CenterlineData:
    """Ghia's centreline velocities for ``reynolds`` in ``REYNOLDS_NUMBERS``."""
    if reynolds not in REYNOLDS_NUMBERS:
        raise ValueError(f"Ghia et al. data available for Re in {REYNOLDS_NUMBERS}; got {reynolds}")
    column = 1 + REYNOLDS_NUMBERS.index(reynolds)
    u_table, v_table = np.array(_TABLE_U), np.array(_TABLE_V)
    return CenterlineData(
        y=_read_only(u_table[:, 0]),
        u=_read_only(u_table[:, column]),
        x=_read_only(v_table[:, 0]),
        v=_read_only(v_table[:, column]),
    )
