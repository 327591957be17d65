"""Builders shared by the test modules (mirrors the reference's tests/helpers.py and the
geometry helpers of tests/test_cuda_backend.py:21-84)."""

import numpy as np
import pytest

from lbm import SimulationParams, make_solver
from lbm.boundaries import channel_walls
from lbm.lattice import LATTICES
from oracle import ReferenceSolver

ALL_LATTICES = pytest.mark.parametrize("vs", list(LATTICES.values()), ids=list(LATTICES))
HIP_BACKENDS = ("hip-pull", "hip-aa", "hip-lds")
HIP_BACKENDS_3D = ("hip-pull", "hip-aa")
EQUIVALENCE_STEPS = 40


def mixed_geometry(nx=45, ny=29):
    """Channel, moving top wall, interior obstacle; not a multiple of any block shape."""
    solid = channel_walls(nx, ny)
    solid[10:16, 12:19] = True
    wall = np.zeros((2, ny, nx))
    wall[0, -1, :] = 0.05
    return solid, wall


def mixed_geometry_3d(nx=24, ny=20, nz=16):
    """Box walls in y and z, periodic x, plate moving in x and z, an obstacle."""
    solid = channel_walls(nx, ny, nz)
    solid[0, :, :] = True
    solid[-1, :, :] = True
    solid[6:10, 7:12, 9:14] = True
    wall = np.zeros((3, nz, ny, nx))
    wall[0, :, -1, :] = 0.05
    wall[2, :, -1, :] = -0.02
    return solid, wall


def mixed_params(dtype, collision, d=2):
    if d == 2:
        solid, wall = mixed_geometry()
        sizes, force = {"nx": 45, "ny": 29}, (1e-5, -2e-6)
    else:
        solid, wall = mixed_geometry_3d()
        sizes, force = {"nx": 24, "ny": 20, "nz": 16}, (1e-5, -2e-6, 3e-6)
    return SimulationParams(
        **sizes, tau=0.7, collision=collision, force=force, solid=solid, wall_velocity=wall, dtype=dtype
    )


def random_state(params, seed=0):
    rng = np.random.default_rng(seed)
    shape = params.shape
    rho = 1 + 0.01 * rng.standard_normal(shape)
    u = 0.02 * rng.standard_normal((params.d, *shape))
    return rho, *u


def solver_for(params, backend, **options):
    if backend == "numpy":
        return ReferenceSolver(params)
    return make_solver(params, backend=backend, **options)


def run_pair(params, backend, steps=EQUIVALENCE_STEPS, seed=0, **options):
    """The NumPy oracle and ``backend`` after the same steps from the same random state."""
    reference = ReferenceSolver(params)
    gpu = solver_for(params, backend, **options)
    for solver in (reference, gpu):
        solver.initialize(*random_state(params, seed))
        solver.run(steps)
    return reference, gpu


def gpu_cases(backends_2d=HIP_BACKENDS, backends_3d=HIP_BACKENDS_3D):
    return [(b, 2) for b in backends_2d] + [(b, 3) for b in backends_3d]


def fluid_diff(params, a, b):
    fluid = params.fluid
    return float(np.abs(a.populations()[:, fluid] - b.populations()[:, fluid]).max())
