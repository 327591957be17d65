"""Validated, immutable description of one simulation (2D or 3D).

The interface is the reference's ``SimulationParams`` (``src/lbm/params.py:86-205``):
the same fields, defaults and validation messages, so a caller's parameter objects
and error handling carry over. The checks themselves are organised here as one
normaliser per field (``_normalise_*``), each returning the stored value or
raising; ``SimulationParams.__post_init__`` only chains them.

Relaxation: nu = cs^2 (tau - 1/2). TRT rates are omega_plus = 1/tau and
omega_minus = 1 / (1/2 + Lambda / (tau - 1/2)); BGK has omega_minus = omega_plus.
Arrays are indexed [y, x] (2D) or [z, y, x] (3D), x fastest.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass, field

import numpy as np

from lbm.ibm import ImmersedBoundary
from lbm.lattice import LATTICES, VelocitySet, default_lattice
from lbm.link_rules import LinkRules

TAU_WARNING = 0.51
COLLISION_MODELS = ("bgk", "trt")
SUPPORTED_DTYPES = ("float32", "float64")


class StabilityWarning(UserWarning):
    """Valid but numerically risky parameters (tau close to 1/2, high Mach number)."""


def check_tau(tau: float) -> None:
    """Reject tau <= 1/2 (no positive viscosity); warn below ``TAU_WARNING``."""
    if not tau > 0.5:
        raise ValueError(f"tau must be > 0.5 (viscosity nu = (tau - 0.5)/3 > 0); got tau={tau}")
    if tau < TAU_WARNING:
        warnings.warn(
            f"tau={tau:.6f} is close to 0.5: BGK becomes unstable in shear layers; "
            "use TRT, refine the grid, or lower the lattice velocity",
            StabilityWarning,
            stacklevel=3,
        )


def _grid_shape(nx: int, ny: int, nz: int | None) -> tuple[int, ...]:
    """(ny, nx) or (nz, ny, nx); every given extent must be positive."""
    named = [("nx", nx), ("ny", ny)] + ([] if nz is None else [("nz", nz)])
    if min(n for _, n in named) <= 0:
        raise ValueError(
            f"{' and '.join(k for k, _ in named)} must be positive; got {', '.join(f'{k}={n}' for k, n in named)}"
        )
    return (ny, nx) if nz is None else (nz, ny, nx)


def _check_choice(value, allowed: tuple[str, ...], what: str) -> None:
    if value not in allowed:
        raise ValueError(f"{what} must be one of {allowed}; got {value!r}")


def _normalise_lattice(spec, d: int) -> VelocitySet:
    """The velocity set named by ``spec`` (default D2Q9 / D3Q19), of dimension ``d``."""
    if spec is None:
        return default_lattice(d)
    lattice = spec if isinstance(spec, VelocitySet) else LATTICES.get(spec)
    if lattice is None:
        raise ValueError(f"unknown lattice {spec!r}; choose from {tuple(LATTICES)}")
    if lattice.d == d:
        return lattice
    raise ValueError(
        f"lattice {lattice.name} is {lattice.d}D but the grid is {d}D (set nz for 3D, leave it unset for 2D)"
    )


def _normalise_solid(mask, shape: tuple[int, ...]) -> np.ndarray:
    """A boolean copy of the mask (all fluid when None)."""
    if mask is None:
        return np.zeros(shape, dtype=bool)
    mask = np.array(mask, dtype=bool, copy=True)
    if mask.shape != shape:
        raise ValueError(f"solid mask must have shape {shape}; got {mask.shape}")
    return mask


def _normalise_walls(velocity, solid: np.ndarray) -> np.ndarray:
    """A float64 copy, shape (d, *shape), kept only at solid nodes (zero when None)."""
    expected = (solid.ndim, *solid.shape)
    if velocity is None:
        return np.zeros(expected)
    velocity = np.asarray(velocity, dtype=np.float64)
    if velocity.shape != expected:
        raise ValueError(f"wall_velocity must have shape {expected}; got {velocity.shape}")
    return np.where(solid, velocity, 0.0)


def _normalise_force(force, d: int) -> tuple[float, ...]:
    components = (0.0,) * d if force is None else tuple(map(float, force))
    if len(components) != d:
        raise ValueError(f"force must have {d} components for a {d}D grid; got {components}")
    return components


def _read_only(array: np.ndarray) -> np.ndarray:
    array.setflags(write=False)
    return array


@dataclass(frozen=True, eq=False)
class SimulationParams:
    """Everything a solver needs to advance one simulation.

    Fields (the reference's): ``nx``, ``ny``, ``tau``, ``collision``
    (``"bgk"``/``"trt"``), ``magic`` (TRT Lambda), ``force`` (uniform body
    force, d components), ``solid`` (bool mask), ``wall_velocity`` (shape
    (d, *shape), read at solid nodes only), ``dtype``, and keyword-only
    ``nz`` (3D when set), ``lattice``, ``link_rules``, ``immersed``.
    """

    nx: int
    ny: int
    tau: float
    collision: str = "bgk"
    magic: float = 3.0 / 16.0
    force: tuple[float, ...] | None = None
    solid: np.ndarray | None = field(default=None, repr=False)
    wall_velocity: np.ndarray | None = field(default=None, repr=False)
    dtype: str = "float32"
    nz: int | None = field(default=None, kw_only=True)
    lattice: str | VelocitySet | None = field(default=None, kw_only=True)
    link_rules: LinkRules | None = field(default=None, kw_only=True, repr=False)
    immersed: ImmersedBoundary | None = field(default=None, kw_only=True, repr=False)

    def __post_init__(self) -> None:
        shape = _grid_shape(self.nx, self.ny, self.nz)
        _check_choice(self.collision, COLLISION_MODELS, "collision")
        if not self.magic > 0:
            raise ValueError(f"TRT magic parameter must be > 0; got magic={self.magic}")
        _check_choice(self.dtype, SUPPORTED_DTYPES, "dtype")
        check_tau(self.tau)
        d = len(shape)
        lattice = _normalise_lattice(self.lattice, d)
        solid = _normalise_solid(self.solid, shape)
        wall = _normalise_walls(self.wall_velocity, solid)
        force = _normalise_force(self.force, d)
        if self.immersed is not None and self.immersed.d != d:
            raise ValueError(f"the immersed body is {self.immersed.d}D but the grid is {d}D")
        if self.link_rules is not None:
            self.link_rules.validate(lattice, solid, wall)
        normalised = {"lattice": lattice, "solid": _read_only(solid), "wall_velocity": _read_only(wall),
                      "force": force, "tau": float(self.tau), "magic": float(self.magic)}
        for name, value in normalised.items():
            object.__setattr__(self, name, value)

    @property
    def d(self) -> int:
        return self.lattice.d

    @property
    def shape(self) -> tuple[int, ...]:
        return self.solid.shape

    @property
    def n_nodes(self) -> int:
        return int(self.solid.size)

    @property
    def fluid(self) -> np.ndarray:
        return ~self.solid

    @property
    def viscosity(self) -> float:
        return self.lattice.cs2 * (self.tau - 0.5)

    @property
    def relaxation_rates(self) -> tuple[float, float]:
        """(omega_plus, omega_minus), in float64 (the reference's params.py:198-205)."""
        omega_plus = 1.0 / self.tau
        if self.collision == "bgk":
            return omega_plus, omega_plus
        return omega_plus, 1.0 / (0.5 + self.magic / (self.tau - 0.5))
