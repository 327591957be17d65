"""Velocity sets: the discrete velocities and weights of the D2Q9 and D3Q19 lattices.

Follows ``src/lbm/lattice.py:52-158`` of the reference (direction numbering,
weights, ``opposite``, ``component``, ``LATTICES``, ``default_lattice``).

The HIP kernels carry the same tables as ``constexpr`` arrays in
``csrc/lbm_lattice.h``; ``tests/test_lattice.py`` reads them back through the
C ABI (``lbm_lattice_tables``) and checks them against this module, so the host
and the device cannot drift apart. The reference generates a CUDA preamble at
run time instead (``lattice.py:165-214``); ahead-of-time compilation makes that
unnecessary here.

D2Q9 numbering (x right, y up)::

    6   2   5
      \\ | /
    3 - 0 - 1
      / | \\
    7   4   8

D3Q19: rest, the six axis links (+x, -x, +y, -y, +z, -z), then the twelve
edge links, each listed right after its opposite.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import cached_property

import numpy as np

CS2 = 1.0 / 3.0
"""Squared lattice speed of sound, shared by both velocity sets."""


def _frozen(values, dtype) -> np.ndarray:
    out = np.array(values, dtype=dtype)
    out.setflags(write=False)
    return out


@dataclass(frozen=True)
class VelocitySet:
    """A DdQq lattice: ``d`` dimensions and ``q = len(c)`` velocities with weights ``w``.

    Loops that build float32 arrays should read the Python tuples ``c`` and
    ``w`` (weak scalars under NumPy 2 promotion); the ``*_array`` views are
    read-only NumPy copies for vectorised host work.
    """

    name: str
    d: int
    c: tuple[tuple[int, ...], ...]
    w: tuple[float, ...]
    cs2: float = CS2

    def __post_init__(self) -> None:
        if len(self.c) != len(self.w):
            raise ValueError(f"{self.name}: {len(self.c)} velocities but {len(self.w)} weights")
        if any(len(v) != self.d for v in self.c):
            raise ValueError(f"{self.name}: every velocity needs {self.d} components")
        if abs(sum(self.w) - 1.0) > 1e-12:
            raise ValueError(f"{self.name}: weights sum to {sum(self.w)}, not 1")
        reversed_set = {tuple(-a for a in v) for v in self.c}
        if reversed_set != set(self.c):
            raise ValueError(f"{self.name}: some velocity has no opposite")

    @property
    def q(self) -> int:
        return len(self.c)

    @cached_property
    def opposite(self) -> tuple[int, ...]:
        """``opposite[i]`` is the index with ``c[opposite[i]] == -c[i]``."""
        index = {v: i for i, v in enumerate(self.c)}
        return tuple(index[tuple(-a for a in v)] for v in self.c)

    def component(self, axis: int) -> tuple[int, ...]:
        """Component ``axis`` (0 = x, 1 = y, 2 = z) of every velocity; zeros past ``d``."""
        if axis >= self.d:
            return (0,) * self.q
        return tuple(v[axis] for v in self.c)

    @cached_property
    def c_array(self) -> np.ndarray:
        return _frozen(self.c, np.int64)

    @cached_property
    def w_array(self) -> np.ndarray:
        return _frozen(self.w, np.float64)

    @property
    def code(self) -> int:
        """Lattice enumerator of the C ABI (``LBM_D2Q9`` = 0, ``LBM_D3Q19`` = 1)."""
        return {"D2Q9": 0, "D3Q19": 1}[self.name]


D2Q9 = VelocitySet(
    name="D2Q9",
    d=2,
    c=((0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)),
    w=(4 / 9,) + (1 / 9,) * 4 + (1 / 36,) * 4,
)

_AXES_3D = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
_EDGES_3D = (
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0),
    (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
    (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
)  # fmt: skip

D3Q19 = VelocitySet(
    name="D3Q19",
    d=3,
    c=((0, 0, 0),) + _AXES_3D + _EDGES_3D,
    w=(1 / 3,) + (1 / 18,) * 6 + (1 / 36,) * 12,
)

LATTICES: dict[str, VelocitySet] = {vs.name: vs for vs in (D2Q9, D3Q19)}


def default_lattice(d: int) -> VelocitySet:
    """D2Q9 for ``d = 2``, D3Q19 for ``d = 3``."""
    table = {2: D2Q9, 3: D3Q19}
    if d not in table:
        raise ValueError(f"no default lattice for dimension {d}; supported: 2, 3")
    return table[d]
