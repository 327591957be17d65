"""Pins the NumPy oracle against the reference's own known-answer tests and closed forms.

Each check mirrors one of the reference's CPU tests (file:line cited) and is chosen so that a
plausible slip in the restatement (a dropped term, a wrong sign or opposite index, a
transposed axis) fails it: moment identities of the equilibrium, conservation and exact
momentum input of the collision, the relaxation rates of each TRT part, Guo source moments,
displacement/wrap of streaming, the bounce-back reflection and Ladd term, uniform flow,
40 F after 40 forced steps, mass conservation in a lid-driven box, and the watchdog.
"""

import numpy as np
import pytest

from lbm import SimulationParams
from lbm.boundaries import cavity_walls, channel_walls
from lbm.lattice import LATTICES
from oracle import reference_solver as ref
from oracle.reference_solver import ReferenceSolver, SimulationDiverged

LATS = pytest.mark.parametrize("name", ["D2Q9", "D3Q19"])
SHAPES = {2: (5, 7), 3: (3, 4, 5)}


def random_macroscopic(lat, seed=0, scale=0.05):
    rng = np.random.default_rng(seed)
    shape = SHAPES[lat.d]
    drho = scale * rng.standard_normal(shape)
    u = tuple(scale * rng.standard_normal(shape) for _ in range(lat.d))
    return drho, u


def random_populations(lat, seed=0, noise=1e-3):
    drho, u = random_macroscopic(lat, seed)
    f = ref.equilibrium(lat, drho, u)
    return f + noise * np.random.default_rng(seed + 1000).standard_normal(f.shape)


def first_moments(lat, f):
    c = np.asarray(lat.c, dtype=float)
    return tuple(np.tensordot(c[:, a], f, axes=1) for a in range(lat.d))


# ---------------------------------------------------------------- lattice (tests/test_lattice.py:23-96)


@LATS
def test_oracle_tables_equal_the_package_tables(name):
    lat, vs = ref.Lattice(name), LATTICES[name]
    assert lat.c == vs.c and lat.w == vs.w and lat.opposite == vs.opposite


def test_d2q9_numbering_known_answer():
    lat = ref.Lattice("D2Q9")
    assert lat.opposite == (0, 3, 4, 1, 2, 7, 8, 5, 6)
    assert lat.c[5] == (1, 1) and lat.c[7] == (-1, -1)
    assert lat.w[0] == 4 / 9 and lat.w[1] == 1 / 9 and lat.w[8] == 1 / 36


def test_d3q19_link_classes_known_answer():
    lat = ref.Lattice("D3Q19")
    norms = [sum(abs(a) for a in c) for c in lat.c]
    assert norms == [0] + [1] * 6 + [2] * 12
    assert lat.opposite == (0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17)


@LATS
def test_moments_are_isotropic_to_fourth_order(name):
    lat = ref.Lattice(name)
    c = np.asarray(lat.c, dtype=float)
    w = np.asarray(lat.w)
    d, cs2 = lat.d, lat.cs2
    eye = np.eye(d)
    assert np.isclose(w.sum(), 1.0)
    np.testing.assert_allclose(np.einsum("i,ia->a", w, c), 0, atol=1e-15)
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", w, c, c), cs2 * eye, atol=1e-15)
    np.testing.assert_allclose(np.einsum("i,ia,ib,ic->abc", w, c, c, c), 0, atol=1e-15)
    fourth = np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c)
    expect = cs2**2 * (
        np.einsum("ab,cd->abcd", eye, eye) + np.einsum("ac,bd->abcd", eye, eye) + np.einsum("ad,bc->abcd", eye, eye)
    )
    np.testing.assert_allclose(fourth, expect, atol=1e-15)


# ---------------------------------------------------------------- equilibrium (tests/test_equilibrium.py:20-58)


@LATS
def test_equilibrium_moments(name):
    lat = ref.Lattice(name)
    drho, u = random_macroscopic(lat)
    feq = ref.equilibrium(lat, drho, u)
    rho = 1 + drho
    np.testing.assert_allclose(feq.sum(0), drho, atol=1e-15)
    for a, j in enumerate(first_moments(lat, feq)):
        np.testing.assert_allclose(j, rho * u[a], atol=1e-15)
    c = np.asarray(lat.c, dtype=float)
    w = np.asarray(lat.w)
    for a in range(lat.d):
        for b in range(lat.d):
            pi = np.tensordot(c[:, a] * c[:, b], feq + w.reshape(-1, *[1] * lat.d), axes=1)
            expect = rho * u[a] * u[b] + (rho / 3.0 if a == b else 0.0)
            np.testing.assert_allclose(pi, expect, atol=1e-14)


def test_equilibrium_at_rest_is_zero_in_shifted_storage():
    lat = ref.Lattice("D3Q19")
    feq = ref.equilibrium(lat, np.zeros((2, 2, 2)), (np.zeros((2, 2, 2)),) * 3)
    assert np.all(feq == 0)


# ---------------------------------------------------------------- collision (tests/test_collision.py:31-99)


@LATS
@pytest.mark.parametrize("collision", [(1.25, 1.25), (1.6, 0.42)])
def test_collision_conserves_mass_and_momentum_without_force(name, collision):
    lat = ref.Lattice(name)
    f = random_populations(lat)
    post = ref.collide(lat, f, *collision, (0.0,) * lat.d)
    np.testing.assert_allclose(post.sum(0), f.sum(0), atol=1e-15)
    for a, b in zip(first_moments(lat, post), first_moments(lat, f), strict=True):
        np.testing.assert_allclose(a, b, atol=1e-15)


@LATS
def test_equilibrium_is_a_fixed_point(name):
    lat = ref.Lattice(name)
    drho, u = random_macroscopic(lat)
    feq = ref.equilibrium(lat, drho, u)
    np.testing.assert_allclose(ref.collide(lat, feq, 1.3, 0.7, (0.0,) * lat.d), feq, atol=1e-16)


@LATS
def test_force_adds_exactly_f_of_momentum(name):
    lat = ref.Lattice(name)
    f = random_populations(lat)
    force = (1e-4, -3e-5, 2e-5)[: lat.d]
    post = ref.collide(lat, f, 1.6, 0.42, force)
    np.testing.assert_allclose(post.sum(0), f.sum(0), atol=1e-15)
    for a, (p, q) in enumerate(zip(first_moments(lat, post), first_moments(lat, f), strict=True)):
        np.testing.assert_allclose(p - q, force[a], atol=1e-15)


@LATS
def test_trt_parts_relax_at_their_own_rates(name):
    lat = ref.Lattice(name)
    f = random_populations(lat)
    wp, wm = 1.6, 0.42
    drho, u = ref.pre_collision_moments(lat, f, (0.0,) * lat.d)
    neq = f - ref.equilibrium(lat, drho, u)
    post = ref.collide(lat, f, wp, wm, (0.0,) * lat.d)
    opp = list(lat.opposite)
    dpost = post - f
    np.testing.assert_allclose((dpost + dpost[opp]) / 2, -wp * (neq + neq[opp]) / 2, atol=1e-16)
    np.testing.assert_allclose((dpost - dpost[opp]) / 2, -wm * (neq - neq[opp]) / 2, atol=1e-16)


@LATS
def test_guo_source_moments(name):
    lat = ref.Lattice(name)
    _, u = random_macroscopic(lat)
    force = (1e-3, 2e-3, -5e-4)[: lat.d]
    wp, wm = 1.4, 0.8
    s = ref.guo_source(lat, u, force, wp, wm)
    np.testing.assert_allclose(s.sum(0), 0, atol=1e-18)
    for a, m in enumerate(first_moments(lat, s)):
        np.testing.assert_allclose(m, (1 - wm / 2) * force[a], atol=1e-18)


# ---------------------------------------------------------------- streaming and walls (tests/test_streaming_boundaries.py)


@LATS
def test_pull_stream_moves_each_population_along_its_velocity_with_wrap(name):
    lat = ref.Lattice(name)
    shape = SHAPES[lat.d]
    for i, c in enumerate(lat.c):
        f = np.zeros((lat.q, *shape))
        f[i][(0,) * lat.d] = 1.0
        g = ref.pull_stream(lat, f)
        target = tuple(np.mod(np.array(c[::-1]), shape))
        assert g[i][target] == 1.0 and g.sum() == 1.0


@LATS
def test_bounce_back_reflects_and_adds_the_ladd_term(name):
    lat = ref.Lattice(name)
    shape = (6, 7) if lat.d == 2 else (5, 6, 7)
    solid = channel_walls(shape[-1], shape[-2], shape[0] if lat.d == 3 else None)
    wall = np.zeros((lat.d, *shape))
    wall[0][solid] = 0.1
    mask, term = ref.bounce_back_links(lat, solid, wall, np.float64)
    for i, (c, w) in enumerate(zip(lat.c, lat.w, strict=True)):
        expected = 6.0 * w * c[0] * 0.1
        np.testing.assert_allclose(term[i][mask[i]], expected, rtol=1e-15)
    f = np.random.default_rng(3).standard_normal((lat.q, *shape))
    streamed = ref.pull_stream(lat, f)
    out = ref.apply_bounce_back(lat, streamed, f, mask, term)
    opp = list(lat.opposite)
    np.testing.assert_array_equal(out[mask], (f[opp] + term)[mask])
    np.testing.assert_array_equal(out[~mask], streamed[~mask])


# ---------------------------------------------------------------- whole steps (tests/test_solver.py:35-132)


def periodic(d=2, n=8, **kw):
    kw.setdefault("dtype", "float64")
    if d == 3:
        kw["nz"] = n
    return SimulationParams(nx=n, ny=n, tau=0.8, **kw)


@pytest.mark.parametrize("d", [2, 3])
def test_initialize_reproduces_requested_fields(d):
    s = ReferenceSolver(periodic(d))
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal((8,) * d)
    u = 0.02 * rng.standard_normal((d, *(8,) * d))
    s.initialize(rho, *u)
    m = s.macroscopic()
    np.testing.assert_allclose(m.rho, rho, atol=1e-14)
    for got, want in zip(m.u, u, strict=True):
        np.testing.assert_allclose(got, want, atol=1e-14)


@pytest.mark.parametrize("d", [2, 3])
def test_uniform_flow_is_an_exact_steady_state(d):
    v = (0.05, -0.02, 0.01)[:d]
    s = ReferenceSolver(periodic(d))
    s.initialize(1.0, *v)
    s.run(50)
    for got, want in zip(s.macroscopic().u, v, strict=True):
        np.testing.assert_allclose(got, want, atol=1e-14)


@pytest.mark.parametrize("d", [2, 3])
def test_body_force_accelerates_periodic_fluid_by_f_per_step(d):
    force = (1e-5, -2e-5, 3e-5)[:d]
    s = ReferenceSolver(periodic(d, force=force))
    s.run(40)
    for got, f in zip(s.macroscopic().u, force, strict=True):
        np.testing.assert_allclose(got, 40 * f, rtol=1e-10)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("collision", ["bgk", "trt"])
def test_closed_cavity_with_moving_lid_conserves_mass(d, collision):
    solid, wall = cavity_walls(n=6, lid_velocity=0.05, dimensions=d)
    kw = {"nz": 8} if d == 3 else {}
    params = SimulationParams(nx=8, ny=8, tau=0.7, collision=collision, solid=solid, wall_velocity=wall,
                              dtype="float64", **kw)
    s = ReferenceSolver(params)
    mass0 = s.macroscopic().total_mass
    s.run(300)
    m = s.macroscopic()
    assert m.total_mass == pytest.approx(mass0, rel=1e-13)
    assert np.abs(m.ux).max() > 1e-3


@pytest.mark.parametrize("d", [2, 3])
def test_non_finite_state_raises_with_location(d):
    s = ReferenceSolver(periodic(d))
    rho = np.ones((8,) * d)
    rho[(3, 5, 2)[:d]] = np.nan
    s.initialize(rho)
    with pytest.raises(SimulationDiverged) as info:
        s.run(10, check_every=5)
    assert info.value.step == 5


@pytest.mark.parametrize("d", [2, 3])
def test_vanishing_density_raises(d):
    s = ReferenceSolver(periodic(d))
    rho = np.ones((8,) * d)
    rho[(slice(1, 4),) * d] = -0.5
    s.initialize(rho)
    with pytest.raises(SimulationDiverged, match="density"):
        s.run(2, check_every=1)


def test_planted_mistake_in_the_collision_is_caught():
    """A sign slip in the antisymmetric relaxation breaks the TRT-rate check above."""
    lat = ref.Lattice("D2Q9")
    f = random_populations(lat)
    wp, wm = 1.6, 0.42
    good = ref.collide(lat, f, wp, wm, (0.0, 0.0))
    drho, u = ref.pre_collision_moments(lat, f, (0.0, 0.0))
    neq = f - ref.equilibrium(lat, drho, u)
    opp = list(lat.opposite)
    bad = f - 0.5 * wp * (neq + neq[opp]) + 0.5 * wm * (neq - neq[opp])
    assert np.abs(good - bad).max() > 1e-6
