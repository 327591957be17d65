"""Batched multi-lattice stepping: every lattice of a batch must equal, bit for bit, the same
lattice run alone through ``HipSolver(params, "pull")`` (no counterpart in the reference)."""

import numpy as np
import pytest

from lbm import SimulationParams, make_solver
from lbm.cases.batched_channels import batched_channels, initial_fields

pytestmark = pytest.mark.gpu


def small_batch(dtype="float32"):
    params = batched_channels(count=5, seed=7, dtype=dtype, h_range=(24, 72))
    params.append(SimulationParams(nx=40, ny=30, tau=0.8, collision="bgk", force=(1e-5, 0.0), dtype=dtype))
    return params


def alone(params, steps, graphs=False):
    solver = make_solver(params, backend="hip-pull", graphs=graphs)
    solver.initialize(*initial_fields(params))
    solver.run(steps)
    return solver.populations()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("steps", [20, 21])
def test_batched_equals_each_lattice_alone(dtype, steps):
    from lbm.batched import BatchedSolver

    params = small_batch(dtype)
    batch = BatchedSolver(params)
    for i, p in enumerate(params):
        batch.initialize(i, *initial_fields(p))
    batch.run(steps)
    for i, p in enumerate(params):
        np.testing.assert_array_equal(batch.populations(i), alone(p, steps))


@pytest.mark.parametrize(("block", "groups"), [((64, 4), 3), ((256, 1), 1000), ((128, 8), 64)])
def test_batched_results_do_not_depend_on_tiles_or_grid(block, groups):
    from lbm.batched import BatchedSolver

    params = small_batch()
    ref = BatchedSolver(params)
    other = BatchedSolver(params, block=block, groups=groups, graphs=False)
    for b in (ref, other):
        for i, p in enumerate(params):
            b.initialize(i, *initial_fields(p))
        b.run(10)
    for i in range(len(params)):
        np.testing.assert_array_equal(other.populations(i), ref.populations(i))


def test_autotuned_batch_still_matches():
    from lbm.batched import BatchedSolver

    params = small_batch()
    batch = BatchedSolver(params, autotune=True)
    for i, p in enumerate(params):
        batch.initialize(i, *initial_fields(p))
    batch.run(8)
    np.testing.assert_array_equal(batch.populations(0), alone(params[0], 8))


def test_graphed_single_solver_equals_eager():
    p = small_batch()[1]
    np.testing.assert_array_equal(alone(p, 13, graphs=True), alone(p, 13, graphs=False))


# runs that capture, then replay, the long and the one-pair graphs, with odd runs between
# them (which swap the buffers) and direct work between the replays
def _graph_runs():
    from lbm.backends.hip_solver import GRAPH_PAIRS

    long = 2 * GRAPH_PAIRS
    return [long + 2, 2 * long + 5, 4, long + 1, 2 * long + 2]


@pytest.mark.parametrize("variant", ["pull", "aa"])
def test_long_graph_replays_of_one_solver_equal_direct_launches(variant):
    p = small_batch()[0]
    runs = _graph_runs()
    if variant == "aa":
        runs = [r + r % 2 for r in runs]
    out = []
    for graphs in (True, False):
        solver = make_solver(p, backend=f"hip-{variant}", graphs=graphs)
        solver.initialize(*initial_fields(p))
        states = []
        for steps in runs:
            solver.run(steps)
            states.append(solver.populations())
        out.append(states)
    for replayed, direct in zip(*out):
        np.testing.assert_array_equal(replayed, direct)


def test_long_graph_replays_of_a_batch_equal_direct_launches():
    from lbm.batched import BatchedSolver

    params = small_batch()
    out = []
    for graphs in (True, False):
        batch = BatchedSolver(params, graphs=graphs)
        for i, p in enumerate(params):
            batch.initialize(i, *initial_fields(p))
        states = []
        for steps in _graph_runs():
            batch.run(steps)
            states.append([batch.populations(i) for i in range(len(params))])
        out.append(states)
    for replayed, direct in zip(*out):
        for a, b in zip(replayed, direct):
            np.testing.assert_array_equal(a, b)


def test_autotuned_solver_keeps_its_state():
    from tests.helpers import mixed_params, random_state

    params = mixed_params("float64", "trt", 3)
    a = make_solver(params, backend="hip-aa")
    b = make_solver(params, backend="hip-aa", autotune=True)
    for s in (a, b):
        s.initialize(*random_state(params))
        s.run(10)
    np.testing.assert_array_equal(a.populations(), b.populations())


def batch_3d(dtype):
    """Uneven D3Q19 lattices: a duct with a Bouzidi sphere and Zou-He ends, a walled box
    with a moving lid, and a periodic box with a body force."""
    from lbm.boundaries import cavity_walls
    from lbm.cases.sphere_drag import duct_sphere

    duct, _, _ = duct_sphere(nx=36, ny=14, nz=12, radius=3.0, centre=(10.0, 7.0, 6.0), dtype=dtype)
    solid, wall = cavity_walls(n=10, lid_velocity=0.04, dimensions=3)
    box = SimulationParams(nx=12, ny=12, nz=12, tau=0.7, collision="trt", solid=solid, wall_velocity=wall,
                           dtype=dtype)
    periodic = SimulationParams(nx=20, ny=9, nz=7, tau=0.8, collision="bgk", force=(1e-5, 0.0, -2e-6), dtype=dtype)
    return [duct, box, periodic]


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_batched_d3q19_equals_each_lattice_alone(dtype):
    """D3Q19 lattices batched (float64 D3Q19 work items are at most 512 lanes; the
    autotuner skips the shapes the kernels were not compiled for)."""
    from lbm.batched import BatchedSolver
    from tests.helpers import random_state

    params = batch_3d(dtype)
    states = [random_state(p, seed=i) for i, p in enumerate(params)]
    for autotune in (False, True):
        batch = BatchedSolver(params, autotune=autotune)
        for i, state in enumerate(states):
            batch.initialize(i, *state)
        batch.run(9)
        for i, (p, state) in enumerate(zip(params, states, strict=True)):
            solver = make_solver(p, backend="hip-pull")
            solver.initialize(*state)
            solver.run(9)
            np.testing.assert_array_equal(batch.populations(i), solver.populations())


def test_stated_batch_of_64_seeded_channels_equals_each_alone_and_the_oracle():
    """The batched configuration as stated: 64 D2Q9 float32 Bouzidi-cylinder channels from
    seed 0 (H in [64, 512], L = 4H, Re in [20, 200], Zou-He inlet and outlet), autotuned
    and graph-captured as the benchmark runs them. After 40 steps every lattice equals the
    same lattice run alone bit for bit, and the smallest equals the NumPy oracle to float32
    tolerance."""
    from lbm.batched import BatchedSolver
    from oracle import ReferenceSolver

    params = batched_channels(count=64, seed=0)
    heights = [p.ny for p in params]
    assert len(params) == 64 and 64 <= min(heights) and max(heights) <= 512
    assert all(p.nx == 4 * p.ny for p in params)
    batch = BatchedSolver(params, autotune=True)
    for i, p in enumerate(params):
        batch.initialize(i, *initial_fields(p))
    batch.run(40)
    for i, p in enumerate(params):
        np.testing.assert_array_equal(batch.populations(i), alone(p, 40), err_msg=f"lattice {i}")
    small = int(np.argmin([p.n_nodes for p in params]))
    p = params[small]
    ref = ReferenceSolver(p)
    ref.initialize(*initial_fields(p))
    ref.run(40)
    fluid = p.fluid
    assert np.abs(batch.populations(small)[:, fluid] - ref.populations()[:, fluid]).max() <= 1e-6
