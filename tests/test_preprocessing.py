"""Vectorised host preprocessing against the node-by-node oracle.

``lbm.open_boundaries.open_channel`` builds Zou-He rows for whole groups of boundary nodes
at once, and ``lbm.geometry.voxelize`` / ``wall_links`` only visit a body's bounding box;
the oracle builds the same things one node (and one term) at a time over the whole grid,
as the reference's builders do (src/lbm/open_boundaries.py:88-186,
src/lbm/geometry.py:136-206). Rows, constants and wall fractions must be bit-identical,
in the same order. The oracle itself is pinned to the reference's golden Zou-He rows.
"""

import numpy as np
import pytest

from lbm.boundaries import channel_walls
from lbm.geometry import circle, sphere, voxelize, wall_links
from lbm.lattice import D2Q9, D3Q19
from lbm.open_boundaries import open_channel
from oracle import preprocessing as ref
from oracle.reference_solver import Lattice
from tests.test_golden import load


def padded_rows(values):
    """Rows of an AffineSet as (direction, node, coefficient) lists without the zero padding."""
    out = []
    for r in range(len(values)):
        keep = values.coeff[r] != 0.0
        out.append(list(zip(values.direction[r][keep].tolist(), values.node[r][keep].tolist(),
                            values.coeff[r][keep].tolist(), strict=True)))
    return out


def same_float(a, b) -> bool:
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def check_open_channel(lattice, solid, inflow, rho_out=1.0):
    got_solid, rules = open_channel(lattice, solid, inflow, rho_out=rho_out)
    exp_solid, nodes, directions, rows, constants = ref.open_channel(Lattice(lattice.name), solid, inflow, rho_out)
    assert np.array_equal(got_solid, exp_solid)
    assert rules.node.tolist() == nodes
    assert rules.direction.tolist() == directions
    got_rows = padded_rows(rules.values)
    assert len(got_rows) == len(rows)
    for r, (g, e) in enumerate(zip(got_rows, rows, strict=True)):
        assert [t[:2] for t in g] == [t[:2] for t in e], f"rule {r}: term order"
        assert same_float([t[2] for t in g], [t[2] for t in e]), f"rule {r}: coefficients"
    assert same_float(rules.values.constant, constants)
    return len(nodes)


def channel_inflow(ny):
    y = np.arange(ny) - 0.5
    h = ny - 2
    return np.where((y > 0) & (y < h), 4 * 0.075 * y * (h - y) / h**2, 0.0)


def test_oracle_reproduces_the_reference_zou_he_rows():
    """The node-by-node builder against rows generated by the reference (2D and D3Q19)."""
    g = load("zou_he")
    for d, lattice in ((2, D2Q9), (3, D3Q19)):
        nx, ny, nz = 12, 9, 5
        walls = channel_walls(nx, ny) if d == 2 else channel_walls(nx, ny, nz)
        inflow = np.zeros((d, *walls.shape[:-1]))
        inflow[0], inflow[1] = 0.04, 0.01
        if d == 3:
            inflow[2] = -0.02
        _, nodes, directions, rows, constants = ref.open_channel(Lattice(lattice.name), walls, inflow, 1.003)
        assert nodes == g[f"node_{d}d"].tolist() and directions == g[f"direction_{d}d"].tolist()
        m = g[f"coeff_{d}d"].shape[1]
        for r, row in enumerate(rows):
            coeff = np.zeros(m)
            coeff[: len(row)] = [t[2] for t in row]
            assert same_float(coeff, g[f"coeff_{d}d"][r])
            assert [t[0] for t in row] == g[f"vdir_{d}d"][r][: len(row)].tolist()
            assert [t[1] for t in row] == g[f"vnode_{d}d"][r][: len(row)].tolist()
        assert same_float(constants, g[f"constant_{d}d"])


def test_2d_channel_with_cylinder_and_parabolic_inflow():
    nx, ny = 60, 26
    solid = channel_walls(nx, ny) | voxelize(circle(14.0, 12.3, 4.0), (ny, nx))
    assert check_open_channel(D2Q9, solid, np.stack([channel_inflow(ny), np.zeros(ny)])) > 0


def test_2d_random_inflow_and_outlet_pressure():
    rng = np.random.default_rng(1)
    nx, ny = 30, 20
    solid = channel_walls(nx, ny)
    solid[7:10, 1:3] = True  # an obstacle against the inlet column
    inflow = np.stack([rng.normal(0.03, 0.02, ny), rng.normal(0, 0.01, ny)])
    check_open_channel(D2Q9, solid, inflow, rho_out=1.004)


def test_2d_periodic_cross_section():
    nx, ny = 30, 16
    solid = voxelize(circle(12.0, 8.0, 3.0), (ny, nx))
    check_open_channel(D2Q9, solid, np.stack([np.full(ny, 0.05), np.zeros(ny)]))


@pytest.mark.parametrize("walls", ["yz", "y", "none"])
def test_3d_duct_with_sphere_and_random_inflow(walls):
    rng = np.random.default_rng(2)
    nx, ny, nz = 16, 11, 9
    solid = np.zeros((nz, ny, nx), dtype=bool)
    if "y" in walls:
        solid[:, 0, :] = solid[:, -1, :] = True
    if "z" in walls:
        solid[0] = solid[-1] = True
    solid |= voxelize(sphere(6.0, 5.0, 4.0, 2.2), solid.shape)
    solid[3:5, 4:6, 1] = True  # obstacle touching the inlet column
    inflow = np.stack([rng.uniform(0, 0.05, (nz, ny)), rng.normal(0, 0.01, (nz, ny)), rng.normal(0, 0.01, (nz, ny))])
    assert check_open_channel(D3Q19, solid, inflow, rho_out=0.998) > 0


def check_links(lattice, shape, body, extra=None):
    vox = voxelize(body, shape)
    assert np.array_equal(vox, ref.voxelize(body, shape))
    solid = vox if extra is None else vox | extra
    links = wall_links(lattice, body, solid)
    node, direction, q = ref.wall_links(Lattice(lattice.name), body, solid)
    assert np.array_equal(links.node, node) and np.array_equal(links.direction, direction)
    assert same_float(links.q, q)
    return len(node)


def test_wall_links_of_a_cylinder_in_a_channel():
    assert check_links(D2Q9, (42, 120), circle(40.3, 20.7, 6.0), channel_walls(120, 42)) > 0


def test_wall_links_across_the_periodic_x_edge_and_against_a_y_edge():
    check_links(D2Q9, (30, 50), circle(1.2, 15.0, 4.5))
    check_links(D2Q9, (30, 50), circle(25.0, 1.5, 3.0))


def test_wall_links_of_spheres():
    check_links(D3Q19, (20, 18, 30), sphere(11.0, 8.5, 9.5, 4.0))
    check_links(D3Q19, (16, 16, 40), sphere(10.0, 7.6, 8.2, 3.3), channel_walls(40, 16, 16))
