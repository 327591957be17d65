"""The on-device divergence watchdog (``lbm_health`` behind ``HipSolver._check_health``)
against the oracle's: the same step, first bad node and reason, on the reference's own
failing cases (tests/test_solver.py:111-132 of the reference: a NaN density, and a 3^d
block of negative density whose centre stays negative after a pull step), for both
layouts (the AA buffer is read in its post-even layout, in place)."""

import numpy as np
import pytest

from lbm import SimulationParams, make_solver
from lbm.boundaries import channel_walls
from lbm.solver import SimulationDiverged
from oracle import ReferenceSolver
from oracle.reference_solver import SimulationDiverged as OracleDiverged

N = 8
CASES = [(b, d, dt) for b in ("hip-pull", "hip-aa") for d in (2, 3) for dt in ("float64", "float32")]


def periodic(d, dtype):
    return SimulationParams(nx=N, ny=N, **({"nz": N} if d == 3 else {}), tau=0.8, dtype=dtype)


def outcome(solver, steps, check_every, error):
    with pytest.raises(error) as info:
        solver.run(steps, check_every=check_every)
    return info.value.step, info.value.location, str(info.value)


@pytest.mark.gpu
@pytest.mark.parametrize(("backend", "d", "dtype"), CASES)
def test_non_finite_density_is_reported_as_the_oracle_reports_it(backend, d, dtype):
    params = periodic(d, dtype)
    rho = np.ones(params.shape)
    rho[(3, 5, 2)[:d]] = np.nan
    ref, gpu = ReferenceSolver(params), make_solver(params, backend=backend)
    for s in (ref, gpu):
        s.initialize(rho)
    expect = outcome(ref, 10, 4, OracleDiverged)
    got = outcome(gpu, 10, 4, SimulationDiverged)
    assert got == expect
    assert got[0] == 4 and "non-finite populations" in got[2] and len(got[1]) == d


@pytest.mark.gpu
@pytest.mark.parametrize(("backend", "d", "dtype"), CASES)
def test_negative_density_block_is_reported_as_the_oracle_reports_it(backend, d, dtype):
    params = periodic(d, dtype)
    rho = np.ones(params.shape)
    rho[(slice(1, 4),) * d] = -0.5
    ref, gpu = ReferenceSolver(params), make_solver(params, backend=backend)
    for s in (ref, gpu):
        s.initialize(rho)
    check = gpu.step_multiple
    expect = outcome(ref, 2, check, OracleDiverged)
    got = outcome(gpu, 2, check, SimulationDiverged)
    assert got == expect
    assert got[0] == check and "non-positive density" in got[2]


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
def test_non_finite_values_inside_a_wall_are_never_read_or_reported(backend):
    """Solid nodes are neither streamed from nor checked: a NaN density planted in a wall
    node leaves the run healthy, on the device as in the oracle."""
    solid = channel_walls(16, 12)
    params = SimulationParams(nx=16, ny=12, tau=0.8, force=(1e-5, 0.0), solid=solid, dtype="float64")
    rho = np.ones(params.shape)
    rho[0, 5] = np.nan
    ref, gpu = ReferenceSolver(params), make_solver(params, backend=backend)
    for s in (ref, gpu):
        s.initialize(rho)
        s.run(20, check_every=2)
    fluid = params.fluid
    np.testing.assert_allclose(gpu.populations()[:, fluid], ref.populations()[:, fluid], rtol=0, atol=1e-15)
