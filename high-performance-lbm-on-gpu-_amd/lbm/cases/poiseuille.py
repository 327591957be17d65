"""Force-driven Poiseuille flow between plates (the case of the reference's ``src/lbm/cases/poiseuille.py``).

Walls at y = 0.5 and ny - 1.5 (halfway bounce-back), width H = ny - 2, exact
steady profile u(y) = F / (2 nu) y (H - y). TRT with Lambda = 3/16 places the
wall exactly, so the error is round-off for any tau; BGK is exact only at
tau = 1/2 + sqrt(3)/4. With ``nz`` the same flow runs as a 3D slab (D3Q19).
"""

from __future__ import annotations

import argparse
import math
from dataclasses import dataclass

import numpy as np

from lbm.backends import make_solver
from lbm.boundaries import channel_walls
from lbm.cases.common import add_backend_arguments, run_until_steady
from lbm.lattice import CS2
from lbm.params import SimulationParams

CHANNEL_LENGTH = 4
BGK_EXACT_TAU = 0.5 + math.sqrt(3.0) / 4.0


@dataclass(frozen=True)
class PoiseuilleResult:
    """Fluid-node distances y from the lower wall, the simulated and exact profiles."""

    y: np.ndarray
    u: np.ndarray
    u_exact: np.ndarray
    steps: int

    @property
    def l2_error(self) -> float:
        """||u - u_exact|| / ||u_exact||."""
        return float(np.linalg.norm(self.u - self.u_exact) / np.linalg.norm(self.u_exact))


def exact_profile(y, height, force, viscosity):
    """u(y) = F / (2 nu) y (H - y) (unit density)."""
    return force / (2.0 * viscosity) * y * (height - y)


def _slab(height: int, tau: float, collision: str, dtype: str, max_velocity: float, nz):
    """Channel parameters (2D, or a 3D slab when nz is set) and the driving force."""
    viscosity = CS2 * (tau - 0.5)
    force = 8.0 * viscosity * max_velocity / height**2
    ny = height + 2
    d = 2 if nz is None else 3
    params = SimulationParams(nx=CHANNEL_LENGTH, ny=ny, nz=nz, tau=tau, collision=collision,
                              force=(force,) + (0.0,) * (d - 1), solid=channel_walls(CHANNEL_LENGTH, ny, nz),
                              dtype=dtype)
    return params, viscosity, force


def _fluid_profile(solver) -> np.ndarray:
    """u_x averaged over every axis but y, at the fluid rows."""
    ux = solver.macroscopic().ux
    y_axis = ux.ndim - 2
    rows = np.take(ux, np.arange(1, ux.shape[y_axis] - 1), axis=y_axis)
    return rows.mean(axis=tuple(a for a in range(ux.ndim) if a != y_axis))


def run_poiseuille(height=16, tau=0.8, collision="trt", backend="hip-aa", dtype="float64", max_velocity=0.01,
                   tol=None, max_steps=500_000, nz=None, solver_factory=None) -> PoiseuilleResult:
    """Run to steady state from rest and compare with the parabola; the force is
    8 nu u_max / H^2. ``solver_factory(params)`` overrides the backend (the test suite
    passes the NumPy oracle). ``tol`` bounds the change of u / u_max per check (default just
    above round-off for the dtype); checks come 2 round(H^2 / (20 nu)) steps apart, a few per
    decay time of the slowest mode."""
    params, viscosity, force = _slab(height, tau, collision, dtype, max_velocity, nz)
    solver = solver_factory(params) if solver_factory else make_solver(params, backend=backend)
    solver.initialize()
    if tol is None:
        tol = {"float64": 1e-12, "float32": 1e-6}[dtype]
    interval = 2 * max(1, round(height**2 / (20.0 * viscosity)))
    steps = run_until_steady(solver, lambda s: _fluid_profile(s) / max_velocity, interval=interval, tol=tol,
                             max_steps=max_steps)
    y = np.arange(height) + 0.5
    return PoiseuilleResult(y=y, u=_fluid_profile(solver), u_exact=exact_profile(y, height, force, viscosity),
                            steps=steps)


def main(argv=None) -> None:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--height", type=int, default=16)
    parser.add_argument("--taus", type=float, nargs="+", default=[0.6, 0.8, 1.0, 1.5, 2.0])
    add_backend_arguments(parser)
    args = parser.parse_args(argv)
    print(f"{'tau':>8} {'BGK L2 error':>14} {'TRT L2 error':>14}")
    for tau in args.taus:
        errors = [run_poiseuille(args.height, tau, c, args.backend, args.dtype).l2_error for c in ("bgk", "trt")]
        print(f"{tau:8.3f} {errors[0]:14.3e} {errors[1]:14.3e}")


if __name__ == "__main__":
    main()
