"""The oracle (and the package's host-side builders) against golden vectors of the reference.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py from the
reference's own NumPy solver. Everything here is rebuilt from this repository alone
(package builders + oracle) and must match bit for bit: the 40-step populations of the
mixed geometries, the Zou-He/Hecht-Harting rule rows (the only part of the D3Q19 open
boundary the invariants cannot fix), Bouzidi rows and the momentum-exchange force, and
the IBM force field, populations and body force.
"""

from pathlib import Path

import numpy as np
import pytest

from lbm import SimulationParams
from lbm.boundaries import channel_walls
from lbm.forces import ForceProbe
from lbm.geometry import circle, sphere, voxelize, wall_links
from lbm.ibm import ImmersedBoundary
from lbm.lattice import default_lattice
from lbm.link_rules import bouzidi_rules
from lbm.open_boundaries import open_channel
from oracle import ReferenceSolver

GOLDEN = Path(__file__).resolve().parent / "golden"
STEPS = 40


def load(name):
    return np.load(GOLDEN / f"{name}.npz", allow_pickle=False)


def random_state(params, seed=0):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal(params.shape)
    u = 0.02 * rng.standard_normal((params.d, *params.shape))
    return rho, *u


def run(params, seed=0, steps=STEPS):
    s = ReferenceSolver(params)
    s.initialize(*random_state(params, seed))
    s.run(steps)
    return s


def mixed(d, collision, dtype):
    g = load("states")
    solid, wall = g[f"solid_{d}d"], g[f"wall_{d}d"]
    nz = None if d == 2 else solid.shape[0]
    force = (1e-5, -2e-6, 3e-6)[:d]
    return SimulationParams(nx=solid.shape[-1], ny=solid.shape[-2], nz=nz, tau=0.7, collision=collision,
                            force=force, solid=solid, wall_velocity=wall, dtype=dtype)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("collision", ["bgk", "trt"])
@pytest.mark.parametrize("d", [2, 3])
def test_mixed_geometry_states_are_bit_identical(d, collision, dtype):
    expected = load("states")[f"f_{d}d_{collision}_{dtype}"]
    got = run(mixed(d, collision, dtype)).populations()
    assert got.dtype == expected.dtype
    np.testing.assert_array_equal(got, expected)


def open_case(d):
    lattice = default_lattice(d)
    nx, ny, nz = 12, 9, 5
    walls = channel_walls(nx, ny) if d == 2 else channel_walls(nx, ny, nz)
    inflow = np.zeros((d, *walls.shape[:-1]))
    inflow[0], inflow[1] = 0.04, 0.01
    if d == 3:
        inflow[2] = -0.02
    solid, rules = open_channel(lattice, walls, inflow, rho_out=1.003)
    params = SimulationParams(nx=nx, ny=ny, nz=nz if d == 3 else None, tau=0.7, collision="trt", solid=solid,
                              dtype="float64", link_rules=rules)
    return params, rules


@pytest.mark.parametrize("d", [2, 3])
def test_zou_he_rule_rows_are_identical(d):
    g = load("zou_he")
    _, rules = open_case(d)
    v = rules.values
    np.testing.assert_array_equal(rules.node, g[f"node_{d}d"])
    np.testing.assert_array_equal(rules.direction, g[f"direction_{d}d"])
    np.testing.assert_array_equal(v.direction, g[f"vdir_{d}d"])
    np.testing.assert_array_equal(v.node, g[f"vnode_{d}d"])
    np.testing.assert_array_equal(v.coeff, g[f"coeff_{d}d"])
    np.testing.assert_array_equal(v.constant, g[f"constant_{d}d"])


@pytest.mark.parametrize("d", [2, 3])
def test_zou_he_states_are_bit_identical(d):
    params, _ = open_case(d)
    np.testing.assert_array_equal(run(params, seed=1).populations(), load("zou_he")[f"f_{d}d"])


def bouzidi_case(d, dtype):
    lattice = default_lattice(d)
    if d == 2:
        nx, ny, nz = 45, 29, None
        solid, body = channel_walls(nx, ny), circle(20.3, 13.6, 5.2)
        wall = np.zeros((2, ny, nx))
        wall[0, -1, :] = 0.05
    else:
        nx, ny, nz = 14, 12, 10
        solid, body = channel_walls(nx, ny, nz), sphere(6.3, 5.6, 4.8, 2.6)
        solid[0] = solid[-1] = True
        wall = np.zeros((3, nz, ny, nx))
        wall[0, :, -1, :] = 0.05
        wall[2, :, -1, :] = -0.02
    force = (1e-5, -2e-6, 3e-6)[:d]
    solid = solid | voxelize(body, solid.shape)
    links = wall_links(lattice, body, solid)
    rules = bouzidi_rules(lattice, links, solid)
    params = SimulationParams(nx=nx, ny=ny, nz=nz, tau=0.7, collision="trt", force=force, solid=solid,
                              wall_velocity=wall, dtype=dtype, link_rules=rules)
    return params, links, rules


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("d", [2, 3])
def test_bouzidi_states_and_force_are_bit_identical(d, dtype):
    g = load("bouzidi")
    params, links, rules = bouzidi_case(d, dtype)
    np.testing.assert_array_equal(links.q, g[f"link_q_{d}d"])
    np.testing.assert_array_equal(rules.values.coeff, g[f"coeff_{d}d"])
    s = run(params)
    np.testing.assert_array_equal(s.populations(), g[f"f_{d}d_{dtype}"])
    np.testing.assert_allclose(ForceProbe.build(params, links).evaluate(s), g[f"force_{d}d_{dtype}"], rtol=1e-13,
                               atol=1e-18)


def ibm_case(d, dtype):
    if d == 2:
        nx, ny, nz = 45, 29, None
        solid, body = channel_walls(45, 29), circle(20.3, 13.6, 5.2)
        wall = np.zeros((2, 29, 45))
        wall[0, -1] = 0.05
    else:
        nx, ny, nz = 16, 14, 12
        solid, body = channel_walls(16, 14, 12), sphere(7.3, 6.6, 5.8, 2.4)
        wall = np.zeros((3, 12, 14, 16))
        wall[0, :, -1] = 0.05
    force = (1e-5, -2e-6, 3e-6)[:d]
    return SimulationParams(nx=nx, ny=ny, nz=nz, tau=0.7, collision="trt", force=force, solid=solid,
                            wall_velocity=wall, dtype=dtype, immersed=ImmersedBoundary.from_ball(body))


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("d", [2, 3])
def test_ibm_field_states_and_body_force_are_bit_identical(d, dtype):
    g = load("ibm")
    params = ibm_case(d, dtype)
    first = run(params, steps=1)
    np.testing.assert_array_equal(first._field, g[f"field1_{d}d_{dtype}"])
    s = run(params)
    np.testing.assert_array_equal(s.populations(), g[f"f_{d}d_{dtype}"])
    np.testing.assert_array_equal(s.immersed_force, g[f"body_force_{d}d_{dtype}"])
