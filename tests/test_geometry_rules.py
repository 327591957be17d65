"""Known answers of the package's host-side geometry, link rules, IBM geometry and lattice
tables (the properties tests/test_geometry_link_rules.py, tests/test_ibm.py and
tests/test_lattice.py of the reference check), each derived from the geometry or the
algebra rather than from another implementation."""

import math

import numpy as np
import pytest

from lbm import SimulationParams
from lbm.boundaries import channel_walls
from lbm.forces import ForceProbe
from lbm.geometry import Pipe, circle, sphere, voxelize, wall_links
from lbm.ibm import ImmersedBoundary, ImmersedForcing, roma
from lbm.lattice import D2Q9, D3Q19, LATTICES, VelocitySet, default_lattice
from lbm.link_rules import AffineSet, LinkRules, bouzidi_rules, compile_affine, compile_rules, reads_other_targets
from lbm.open_boundaries import open_channel

# ---------------------------------------------------------------- shapes


def test_ball_entry_fractions():
    disc = circle(1.0, 1.0, 2.0)
    # from (4, 1) towards -x the disc starts at x = 3: one unit along a unit step
    np.testing.assert_allclose(disc.ray_fraction([[4.0, 1.0]], [[-1.0, 0.0]]), [1.0])
    # diagonal step from (3, 3): |(3 - t - 1, 3 - t - 1)| = 2  =>  t = 2 - sqrt(2)
    np.testing.assert_allclose(disc.ray_fraction([[3.0, 3.0]], [[-1.0, -1.0]]), [2.0 - math.sqrt(2.0)])
    ball = sphere(0.0, 0.0, 0.0, 2.0)
    np.testing.assert_allclose(ball.ray_fraction([[0.0, 0.0, 2.5]], [[0.0, 0.0, -1.0]]), [0.5])
    assert ball.measure == pytest.approx(4 / 3 * math.pi * 8)


def test_pipe_exit_fraction_and_inside_test():
    pipe = Pipe((0.0, 0.0), 2.0)  # solid outside radius 2 around the x axis
    # from (0, 1.5, 0) stepping +y the wall is at y = 2: half a step
    np.testing.assert_allclose(pipe.ray_fraction([[0.0, 1.5, 0.0]], [[0.0, 1.0, 0.0]]), [0.5])
    assert pipe.contains(np.array([7.0, 0.0, 2.5])) and not pipe.contains(np.array([7.0, 1.0, 1.0]))


def test_voxelize_counts_nodes_on_and_inside_the_surface():
    # radius 1 around a node: the node and its 4 axis neighbours (at distance exactly 1)
    mask = voxelize(circle(5.0, 4.0, 1.0), (9, 11))
    assert mask.sum() == 5 and mask[4, 5] and mask[3, 5] and mask[4, 6] and not mask[3, 4]
    assert voxelize(sphere(4.0, 4.0, 4.0, 1.0), (9, 9, 9)).sum() == 7


@pytest.mark.parametrize("body", [circle(10.3, 8.6, 4.2), sphere(8.4, 7.1, 6.6, 3.7)])
def test_wall_links_run_from_fluid_into_the_body_with_the_wall_at_q(body):
    lattice = default_lattice(body.d)
    shape = (17, 21) if body.d == 2 else (14, 15, 17)
    solid = voxelize(body, shape)
    links = wall_links(lattice, body, solid)
    assert len(links) > 0 and not solid.ravel()[links.node].any()
    assert np.all((links.q > 0) & (links.q <= 1))
    x = np.stack(np.unravel_index(links.node, shape)[::-1], axis=-1).astype(float)
    c = lattice.c_array[links.direction]
    assert body.contains(x - c).all()  # the link ends inside the body
    np.testing.assert_allclose(np.linalg.norm(x - links.q[:, None] * c - body.center, axis=1), body.radius)
    if body.d == 3:  # every one of the 18 neighbour directions meets a sphere this size
        assert set(links.direction.tolist()) == set(range(1, 19))


def test_wall_links_need_the_body_marked_solid():
    with pytest.raises(ValueError, match="solid"):
        wall_links(D2Q9, circle(5.0, 5.0, 2.0), np.zeros((11, 11), dtype=bool))


# ---------------------------------------------------------------- link rules


def disc_channel(nx=33, ny=21):
    body = circle(15.3, 10.4, 4.6)
    solid = channel_walls(nx, ny) | voxelize(body, (ny, nx))
    return body, solid, wall_links(D2Q9, body, solid)


def test_bouzidi_coefficients_interpolate_linearly_and_sum_to_one():
    _, solid, links = disc_channel()
    rules = bouzidi_rules(D2Q9, links, solid)
    coeff = rules.values.coeff
    np.testing.assert_allclose(coeff.sum(axis=1), 1.0)
    q = {(x, j): qq for x, j, qq in zip(links.node.tolist(), links.direction.tolist(), links.q, strict=True)}
    for r, key in enumerate(zip(rules.node.tolist(), rules.direction.tolist(), strict=True)):
        qq = q[key]
        if coeff[r, 1] != 0.0:  # a two-term interpolation (else the plain bounce-back fallback)
            assert coeff[r, 0] == pytest.approx(2 * qq if qq < 0.5 else 0.5 / qq)


def test_bouzidi_at_q_one_half_is_plain_bounce_back():
    # a flat wall half-way between nodes: y = 2.5 splits rows 2 and 3 of a channel
    nx, ny = 12, 8
    solid = np.zeros((ny, nx), dtype=bool)
    solid[:3] = True
    solid[-1] = True  # a top wall that is not part of the body (no links across the periodic edge)

    class Half:
        d = 2

        def contains(self, p):
            return np.asarray(p)[..., 1] <= 2.5

        def ray_fraction(self, origin, direction):
            o, d = np.asarray(origin, dtype=float), np.asarray(direction, dtype=float)
            return (o[..., 1] - 2.5) / -d[..., 1]

    links = wall_links(D2Q9, Half(), solid)
    np.testing.assert_allclose(links.q, 0.5)
    rules = bouzidi_rules(D2Q9, links, solid)
    np.testing.assert_allclose(rules.values.coeff[:, 0], 1.0)  # f_j = f*_i(x): bounce-back
    assert np.all(rules.values.coeff[:, 1:] == 0.0)


def test_rules_must_sit_on_fluid_to_resting_solid_links():
    _, solid, links = disc_channel()
    rules = bouzidi_rules(D2Q9, links, solid)
    ny, nx = solid.shape
    one = AffineSet.from_lists([[(0, 0, 1.0)]])
    with pytest.raises(ValueError, match="fluid nodes"):
        LinkRules(np.array([0]), np.array([2]), one).validate(D2Q9, solid, np.zeros((2, ny, nx)))
    fluid_node = int(np.ravel_multi_index((10, 3), solid.shape))  # upstream along +x is fluid
    with pytest.raises(ValueError, match="does not start at a solid node"):
        LinkRules(np.array([fluid_node]), np.array([1]), one).validate(D2Q9, solid, np.zeros((2, ny, nx)))
    moving = np.zeros((2, ny, nx))
    moving[0][solid] = 0.01
    with pytest.raises(ValueError, match="at rest"):
        rules.validate(D2Q9, solid, moving)
    with pytest.raises(ValueError, match="more than one rule"):
        LinkRules.concat(rules, rules)


def test_rules_reading_another_rules_target_are_detected():
    target = np.array([10, 20])
    coeff = np.array([[0.5, 0.5], [1.0, 0.0]])
    # rule 0 reads slot 20, which rule 1 writes: the two-phase pass is needed
    assert reads_other_targets(target, np.array([[20, 3], [4, 0]]), coeff)
    # a rule reading its own target is no dependency (its lane reads before it writes)
    assert not reads_other_targets(target, np.array([[10, 3], [4, 0]]), coeff)
    # a zero coefficient is padding, its slot never read
    assert not reads_other_targets(target, np.array([[3, 20], [4, 0]]), np.array([[1.0, 0.0], [1.0, 0.0]]))
    assert not reads_other_targets(np.zeros(0, np.int64), np.zeros((0, 2), np.int64), np.zeros((0, 2)))


def test_body_and_open_end_rules_are_one_launch_safe_in_every_layout():
    body, solid, links = disc_channel()
    ny, nx = solid.shape
    y = np.arange(ny) - 0.5
    inflow = np.stack([np.where((y > 0) & (y < ny - 2), 0.01, 0.0), np.zeros(ny)])
    solid, open_rules = open_channel(D2Q9, solid, inflow)
    links = wall_links(D2Q9, body, solid)
    rules = LinkRules.concat(open_rules, bouzidi_rules(D2Q9, links, solid))
    walls = np.zeros((2, ny, nx))
    for layout in ("pull", "aa_odd", "aa_even"):
        c = compile_rules(D2Q9, rules, layout, solid, walls)
        assert not reads_other_targets(c.target, c.values.slots, c.values.coeff), layout
    # in the pull layout a Bouzidi row with q < 1/2 reads f*_opp(j)(x), the slot it writes
    c = compile_rules(D2Q9, rules, "pull", solid, walls)
    assert ((c.values.slots == c.target[:, None]) & (c.values.coeff != 0)).any()


def test_unknown_layout_is_rejected():
    _, solid, links = disc_channel()
    values = bouzidi_rules(D2Q9, links, solid).values
    with pytest.raises(ValueError, match="layout"):
        compile_affine(D2Q9, values, "zigzag", solid, np.zeros((2, *solid.shape)))


def test_probe_refuses_a_solver_of_other_params():
    body, solid, links = disc_channel()
    params = SimulationParams(nx=33, ny=21, tau=0.8, solid=solid)
    probe = ForceProbe.build(params, links)

    class Other:
        params = SimulationParams(nx=33, ny=21, tau=0.8, solid=solid)

    with pytest.raises(ValueError, match="different SimulationParams"):
        probe.evaluate(Other())
    assert probe.weight.shape == (len(links), 2)


def test_open_channel_needs_room_for_both_layers():
    with pytest.raises(ValueError, match="4 nodes"):
        open_channel(D2Q9, np.zeros((6, 3), dtype=bool), np.zeros((2, 6)))


# ---------------------------------------------------------------- IBM geometry


def test_roma_kernel_is_a_partition_of_unity_with_zero_first_moment():
    x = np.random.default_rng(0).uniform(0, 1, 50)
    k = np.arange(-3, 4)
    w = roma(x[:, None] - k[None, :])
    np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose((w * (x[:, None] - k)).sum(axis=1), 0.0, atol=1e-14)
    assert roma(np.array([1.5, -1.5, 2.0])).tolist() == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("body", [circle(20.0, 15.0, 6.0), sphere(12.0, 11.0, 10.0, 4.0)])
def test_markers_cover_the_surface_and_stencil_weights_sum_to_one(body):
    ib = ImmersedBoundary.from_ball(body)
    np.testing.assert_allclose(np.linalg.norm(ib.markers - body.center, axis=1), body.radius)
    surface = 2 * math.pi * body.radius if body.d == 2 else 4 * math.pi * body.radius**2
    assert ib.areas.sum() == pytest.approx(surface)
    shape = (30, 40) if body.d == 2 else (22, 24, 26)
    _, weights = ib.stencil(shape)
    np.testing.assert_allclose(weights.sum(axis=1), 1.0, atol=1e-14)


def test_ibm_stencils_must_stay_inside_and_clear_of_walls():
    near_edge = ImmersedBoundary.from_ball(circle(1.6, 15.0, 1.5))  # a stencil reaches x = -1
    with pytest.raises(ValueError, match="inside the domain"):
        near_edge.stencil((30, 40))
    walls = channel_walls(40, 30)
    body = ImmersedBoundary.from_ball(circle(20.0, 4.0, 2.0))
    with pytest.raises(ValueError, match="solid"):
        ImmersedForcing(SimulationParams(nx=40, ny=30, tau=0.8, solid=walls, immersed=body))
    with pytest.raises(ValueError, match="3D"):
        SimulationParams(nx=40, ny=30, tau=0.8, immersed=ImmersedBoundary.from_ball(sphere(9, 9, 9, 3)))


def test_pull_moment_rows_are_sums_of_the_pulled_populations():
    """The one-launch forward step sums each stencil node's moment rows from the populations
    its lane pulls. The compiled pull rows are checked here by coordinates with periodic wrap:
    row a n + k has, in lattice order, the directions i with weight w_a,i != 0 (1 for the
    density, c_i,a for momentum a) on slot (i, x_k - c_i), then zero coefficients, constant
    0. A changed coefficient, slot, padding term or constant is refused."""
    from dataclasses import replace

    from lbm.ibm import rows_are_pulled_moments

    params = SimulationParams(nx=40, ny=30, tau=0.8, solid=channel_walls(40, 30),
                              immersed=ImmersedBoundary.from_ball(circle(17.3, 14.6, 4.2)))
    ibm = ImmersedForcing(params)
    rows = ibm.compiled("pull")
    nn, (ny, nx) = ibm.n_nodes, params.shape
    for k, node in enumerate(ibm.nodes):
        y, x = divmod(int(node), nx)
        for a, weight in enumerate(([1] * 9, [c[0] for c in D2Q9.c], [c[1] for c in D2Q9.c])):
            used = [i for i in range(9) if weight[i] != 0]
            r = a * nn + k
            for t, i in enumerate(used):
                cx, cy = D2Q9.c[i]
                assert rows.slots[r, t] == i * nx * ny + ((y - cy) % ny) * nx + (x - cx) % nx
                assert rows.coeff[r, t] == weight[i]
            assert np.all(rows.coeff[r, len(used):] == 0) and rows.constant[r] == 0
    assert rows_are_pulled_moments(D2Q9, ibm.nodes, params.shape, rows)
    for field, index, value in (("coeff", (nn, 0), 2.0), ("slots", (3, 4), 0), ("coeff", (2 * nn + 1, 8), 1.0),
                                ("constant", (5,), 1e-3)):
        planted = np.array(getattr(rows, field), copy=True)
        planted[index] = value
        assert not rows_are_pulled_moments(D2Q9, ibm.nodes, params.shape, replace(rows, **{field: planted}))


# ---------------------------------------------------------------- lattices


@pytest.mark.parametrize("vs", [D2Q9, D3Q19])
def test_velocity_set_tables(vs):
    c = vs.c_array
    assert vs.q == len(c) and np.isclose(vs.w_array.sum(), 1.0)
    opp = np.asarray(vs.opposite)
    assert np.all(c[opp] == -c) and np.all(opp[opp] == np.arange(vs.q))
    assert not vs.c_array.flags.writeable and not vs.w_array.flags.writeable
    assert LATTICES[vs.name] is vs and default_lattice(vs.d) is vs and hash(vs) == hash(vs)
    assert vs.component(2) == ((0,) * vs.q if vs.d == 2 else tuple(c[:, 2]))


def test_inconsistent_velocity_sets_are_rejected():
    with pytest.raises(ValueError, match="weights"):
        VelocitySet("X", 1, ((0,), (1,), (-1,)), (0.5, 0.5))
    with pytest.raises(ValueError, match="sum"):
        VelocitySet("X", 1, ((0,), (1,), (-1,)), (0.5, 0.3, 0.3))
    with pytest.raises(ValueError, match="opposite"):
        VelocitySet("X", 1, ((0,), (1,), (2,)), (0.5, 0.25, 0.25))
    with pytest.raises(ValueError, match="components"):
        VelocitySet("X", 2, ((0, 0), (1,)), (0.5, 0.5))
    with pytest.raises(ValueError, match="dimension"):
        default_lattice(4)


def test_link_rule_index_orders_rules_by_node_and_direction():
    """The rule-node list of the in-step rule evaluation: rows sorted by (node, direction),
    one entry per node with rows, node | direction bits << 32 and the first row, for a rule
    set given in scrambled order."""
    from lbm.backends.hip_solver import link_rule_index

    solid = np.ones((5, 10), dtype=bool)  # n = 50 nodes; every upstream node is solid
    n = solid.size
    opp = np.asarray(D2Q9.opposite)
    links = [(7, 1), (7, 5), (7, 8), (3, 2), (40, 6), (3, 4)]  # (node, direction j)
    target = np.array([opp[j] * n + x for x, j in links], dtype=np.int64)
    order, info, first = link_rule_index(D2Q9, target[::-1], solid)
    assert [links[::-1][r] for r in order] == sorted(links)
    every = sum(1 << i for i in range(1, 9))
    assert info.tolist() == [[3, every, (1 << 2) | (1 << 4), 0], [7, every, (1 << 1) | (1 << 5) | (1 << 8), 2],
                             [40, every, 1 << 6, 5]]
    assert first.tolist() == [0, 2, 5, 6]
    with pytest.raises(ValueError, match="same slot"):
        link_rule_index(D2Q9, np.r_[target, target[:1]], solid)


def test_link_rule_index_records_the_solid_upstream_directions():
    """The blocked set of a rule node: bit i when x - c_i is solid, with periodic wrap."""
    from lbm.backends.hip_solver import link_rule_index

    solid = np.zeros((6, 8), dtype=bool)
    solid[:, 0] = True  # the first column: upstream of direction 1 (c = (1, 0)) at x = 1
    node = 2 * 8 + 1  # (y, x) = (2, 1)
    opp = np.asarray(D2Q9.opposite)
    _, info, _ = link_rule_index(D2Q9, np.array([opp[1] * solid.size + node]), solid)
    blocked = [i for i, c in enumerate(D2Q9.c) if any(c) and solid[(2 - c[1]) % 6, (1 - c[0]) % 8]]
    assert blocked == [1, 5, 8]
    assert info[0].tolist() == [node, sum(1 << i for i in blocked), 1 << 1, 0]
    with pytest.raises(ValueError, match="not solid"):
        link_rule_index(D2Q9, np.array([opp[2] * solid.size + node]), solid)
