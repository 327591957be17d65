"""CPU oracle: NumPy restatement of the reference solver's time step.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU baseline); the ``lbm`` package never imports it.

It restates, as whole-array NumPy expressions written in the same order as
the reference (ShauryaAttreya37/High-Performance-LBM-On-GPU-), so that every
population rounds exactly as the reference's NumPy backend does:

* lattice tables ........................ src/lbm/lattice.py:114-146
* equilibrium ........................... src/lbm/equilibrium.py:31-58
* momentum / moments / collide .......... src/lbm/collision.py:96-151
* Guo source ............................ src/lbm/forcing.py:37-56
* pull streaming ........................ src/lbm/streaming.py:27-37
* bounce-back links / application ....... src/lbm/boundaries.py:53-78 (as file lines 282-322)
* affine evaluation, reference layout ... src/lbm/link_rules.py:219-220, 243-272
* IBM forcing (Roma kernel, multi-direct) src/lbm/ibm.py:53-58, 107-120, 159-214
* ReferenceSolver / run / health ........ src/lbm/solver.py:101-172, 224-277
* macroscopic from post-collision ....... src/lbm/diagnostics.py:72-80

Inputs are the package's ``SimulationParams`` (the run description); the
oracle resolves everything else (lattice constants, rule slots in the
reference layout, IBM stencils and moment rows) on its own.

Parity status: every function here is pinned by tests/test_oracle.py against
the reference's known-answer tests and closed forms (conservation, exact
momentum input, Poiseuille, Taylor-Green/ABC order, Zou-He exact imposition,
Bouzidi q = 1/2, force balance, Roma partition of unity) and, where the
reference's own tests cannot pin the exact bits, against golden vectors
generated from the reference in tests/golden/.
"""

from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------- lattices

_C = {
    "D2Q9": ((0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)),
    "D3Q19": (
        (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
        (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
        (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
    ),
}  # fmt: skip
_W = {
    "D2Q9": (4 / 9,) + (1 / 9,) * 4 + (1 / 36,) * 4,
    "D3Q19": (1 / 3,) + (1 / 18,) * 6 + (1 / 36,) * 12,
}
CS2 = 1.0 / 3.0


class Lattice:
    def __init__(self, name: str):
        self.name = name
        self.c = _C[name]
        self.w = _W[name]
        self.d = len(self.c[0])
        self.q = len(self.c)
        self.cs2 = CS2
        self.opposite = tuple(self.c.index(tuple(-a for a in v)) for v in self.c)

    def component(self, axis: int):
        return tuple(v[axis] for v in self.c) if axis < self.d else (0,) * self.q


def lattice_for(params) -> Lattice:
    return Lattice(params.lattice.name)


# ---------------------------------------------------------------- physics


def dot(c, vector):
    """c . v over the non-zero components of c; the rest direction gives the int 0."""
    return sum(c_a * v_a for c_a, v_a in zip(c, vector, strict=True) if c_a)


def equilibrium(lat: Lattice, delta_rho, u):
    rho = 1.0 + delta_rho
    usq = 1.5 * sum(u_a * u_a for u_a in u)
    terms = []
    for c_i, w_i in zip(lat.c, lat.w, strict=True):
        cu = 3.0 * dot(c_i, u)
        terms.append(w_i * (delta_rho + rho * (cu + 0.5 * cu * cu - usq)))
    return np.stack(terms)


def momentum(lat: Lattice, f):
    return tuple(
        sum(c_a * f[i] for i, c_a in enumerate(lat.component(axis)) if c_a) for axis in range(lat.d)
    )


def pre_collision_moments(lat: Lattice, f, force):
    delta_rho = f.sum(axis=0)
    rho = 1.0 + delta_rho
    u = tuple((j_a + 0.5 * f_a) / rho for j_a, f_a in zip(momentum(lat, f), force, strict=True))
    return delta_rho, u


def guo_source(lat: Lattice, u, force, omega_plus, omega_minus):
    u_dot_f = sum(u_a * f_a for u_a, f_a in zip(u, force, strict=True))
    a_plus = 1.0 - 0.5 * omega_plus
    a_minus = 1.0 - 0.5 * omega_minus
    terms = []
    for c_i, w_i in zip(lat.c, lat.w, strict=True):
        c_dot_f = dot(c_i, force)
        c_dot_u = dot(c_i, u)
        symmetric = 9.0 * c_dot_u * c_dot_f - 3.0 * u_dot_f
        antisymmetric = 3.0 * c_dot_f
        terms.append(w_i * (a_plus * symmetric + a_minus * antisymmetric))
    return np.stack(terms)


def _is_zero(value) -> bool:
    return isinstance(value, int | float) and value == 0.0


def collide(lat: Lattice, f, omega_plus, omega_minus, force):
    delta_rho, u = pre_collision_moments(lat, f, force)
    non_eq = f - equilibrium(lat, delta_rho, u)
    non_eq_opp = non_eq[list(lat.opposite)]
    post = f - (0.5 * omega_plus) * (non_eq + non_eq_opp) - (0.5 * omega_minus) * (non_eq - non_eq_opp)
    if any(not _is_zero(f_a) for f_a in force):
        post = post + guo_source(lat, u, force, omega_plus, omega_minus)
    return post


def pull_stream(lat: Lattice, f_post):
    axes = tuple(range(lat.d))
    return np.stack([np.roll(f_post[i], shift=tuple(reversed(c)), axis=axes) for i, c in enumerate(lat.c)])


def bounce_back_links(lat: Lattice, solid, wall_velocity, dtype):
    solid = np.asarray(solid, dtype=bool)
    fluid = ~solid
    walls = [np.asarray(wall_velocity[a], dtype=dtype) for a in range(lat.d)]
    axes = tuple(range(lat.d))
    masks, terms = [], []
    for c_i, w_i in zip(lat.c, lat.w, strict=True):
        shift = tuple(reversed(c_i))
        from_solid = np.roll(solid, shift, axis=axes) & fluid
        c_dot_uw = sum(c_a * np.roll(walls[a], shift, axis=axes) for a, c_a in enumerate(c_i) if c_a)
        term = (2.0 * w_i / lat.cs2) * c_dot_uw
        masks.append(from_solid)
        terms.append(np.where(from_solid, term, 0.0).astype(dtype))
    return np.stack(masks), np.stack(terms)


def apply_bounce_back(lat: Lattice, streamed, f_post, mask, wall_term):
    reflected = f_post[list(lat.opposite)] + wall_term
    return np.where(mask, reflected, streamed)


def evaluate_affine(slots, coeff, constant, flat):
    gathered = flat[slots]
    return constant + (coeff * gathered).sum(axis=1)


def macroscopic_from_post_collision(lat: Lattice, f_post, force, solid):
    rho = 1.0 + f_post.sum(axis=0)
    u = tuple((j_a - 0.5 * f_a) / rho for j_a, f_a in zip(momentum(lat, f_post), force, strict=True))
    return np.where(solid, 1.0, rho), tuple(np.where(solid, 0.0, u_a) for u_a in u)


def neighbour(nodes, c, shape):
    coords = np.unravel_index(np.asarray(nodes, dtype=np.int64), shape)
    moved = [(x + s) % n for x, s, n in zip(coords, tuple(c)[::-1], shape, strict=True)]
    return np.ravel_multi_index(moved, shape)


# ---------------------------------------------------------------- IBM


def roma(r):
    a = np.abs(np.asarray(r, dtype=np.float64))
    inner = (1.0 + np.sqrt(np.maximum(1.0 - 3.0 * a * a, 0.0))) / 3.0
    outer = (5.0 - 3.0 * a - np.sqrt(np.maximum(1.0 - 3.0 * (1.0 - a) ** 2, 0.0))) / 6.0
    return np.where(a <= 0.5, inner, np.where(a <= 1.5, outer, 0.0))


class ImmersedForcing:
    """Restatement of ``ImmersedForcing`` for the reference layout."""

    def __init__(self, params, lat: Lattice):
        body = params.immersed
        shape = params.shape
        d = lat.d
        markers = np.asarray(body.markers, dtype=np.float64)
        base = np.floor(markers - 0.5).astype(np.int64) + 0
        offsets = np.stack(np.meshgrid(*(np.arange(3),) * d, indexing="ij"), axis=-1).reshape(-1, d)
        points = base[:, None, :] + offsets[None, :, :]
        weights = np.prod(roma(points - markers[:, None, :]), axis=-1)
        if np.any(points < 0) or np.any(points >= np.asarray(shape[::-1])):
            raise ValueError("every marker must sit at least 2 nodes inside the domain")
        stencil = np.ravel_multi_index(tuple(points[..., a] for a in range(d))[::-1], shape)
        nodes, inverse = np.unique(stencil, return_inverse=True)
        flat_solid = np.asarray(params.solid, dtype=bool).ravel()
        for c in lat.c:
            if np.any(flat_solid[neighbour(nodes, c, shape)]):
                raise ValueError("immersed-boundary stencils must not touch solid nodes")
        self.nodes = nodes
        self.inverse = inverse.reshape(stencil.shape)
        dtype = params.dtype
        self.weights = weights.astype(dtype)
        self.areas = np.asarray(body.areas).astype(dtype)
        self.d = d
        self.n_nodes = len(nodes)
        self.iterations = body.iterations
        self.uniform_force = tuple(float(f) for f in params.force)
        self.body_force = np.zeros(d, dtype=np.float64)
        # Moment rows (rho - 1, then rho u per axis) over f*_k(x - c_k), padded to q terms.
        sources = [neighbour(nodes, [-a for a in c_k], shape) for c_k in lat.c]
        n_rows = (d + 1) * len(nodes)
        slots = np.zeros((n_rows, lat.q), dtype=np.int64)
        coeff = np.zeros((n_rows, lat.q))
        n_total = int(np.prod(shape))
        row = 0
        for component in range(d + 1):
            for n in range(len(nodes)):
                t = 0
                for k, c_k in enumerate(lat.c):
                    a = 1.0 if component == 0 else float(c_k[component - 1])
                    if a:
                        slots[row, t] = k * n_total + sources[k][n]
                        coeff[row, t] = a
                        t += 1
                row += 1
        self.slots = slots
        self.coeff = coeff.astype(dtype)
        self.constant = np.zeros(n_rows, dtype=dtype)

    def force_field(self, state_flat):
        moments = evaluate_affine(self.slots, self.coeff, self.constant, state_flat).reshape(self.d + 1, self.n_nodes)
        rho = 1.0 + moments[0]
        u_star = np.stack([(moments[1 + a] + 0.5 * self.uniform_force[a]) / rho for a in range(self.d)])
        rho_marker = (self.weights * rho[self.inverse]).sum(axis=1)
        force = np.zeros((self.d, self.n_nodes), dtype=u_star.dtype)
        total = np.zeros(self.d, dtype=u_star.dtype)
        flat_inverse = self.inverse.reshape(-1)
        for _ in range(self.iterations):
            u = u_star + force / (2.0 * rho)
            u_marker = (self.weights[None] * u[:, self.inverse]).sum(axis=2)
            f_marker = 2.0 * rho_marker * (0.0 - u_marker) * self.areas
            total = total + f_marker.sum(axis=1)
            spread = (self.weights[None] * f_marker[:, :, None]).reshape(self.d, -1)
            for a in range(self.d):
                force[a] = force[a] + np.bincount(flat_inverse, weights=spread[a], minlength=self.n_nodes)
        self.body_force = -total
        return force


# ---------------------------------------------------------------- solver


class SimulationDiverged(RuntimeError):
    """The reference's exception and message (src/lbm/solver.py:39-52)."""

    def __init__(self, step, location, reason):
        self.step = step
        self.location = location
        names = "(y, x)" if len(location) == 2 else "(z, y, x)"
        super().__init__(f"simulation diverged by step {step}: {reason} at node {names}={location}")


class Fields:
    """rho, u and the solid mask of one snapshot (the oracle's MacroscopicFields)."""

    def __init__(self, rho, u, solid):
        self.rho = rho
        self.u = tuple(u)
        self.solid = solid

    @property
    def ux(self):
        return self.u[0]

    @property
    def uy(self):
        return self.u[1]

    @property
    def uz(self):
        return self.u[2]

    @property
    def total_mass(self) -> float:
        return float(self.rho[~self.solid].sum(dtype=np.float64))


class ReferenceSolver:
    """Whole-array NumPy solver: the oracle every HIP result is compared with."""

    step_multiple = 1
    state_layout = "reference"

    def __init__(self, params):
        self.params = params
        self.lat = lattice_for(params)
        self.time_step = 0
        self._fluid = np.asarray(params.fluid)
        self._mask, self._wall_term = bounce_back_links(
            self.lat, params.solid, params.wall_velocity, np.dtype(params.dtype)
        )
        self._rates = params.relaxation_rates
        self._rules = None
        rules = params.link_rules
        if rules is not None and len(rules):
            n = params.n_nodes
            values = rules.values
            self._rules = (
                np.asarray(rules.direction) * n + np.asarray(rules.node),
                values.direction * n + values.node,
                values.coeff.astype(params.dtype),
                values.constant.astype(params.dtype),
            )
        self._ibm = self._field = None
        if params.immersed is not None:
            self._ibm = ImmersedForcing(params, self.lat)
            self._field = np.zeros((params.d, *params.shape), dtype=params.dtype)
        self.initialize()

    def _velocity_components(self, ux, uy, uz):
        if self.params.d == 3:
            return (ux, uy, uz)
        if np.any(np.asarray(uz) != 0):
            raise ValueError("a 2D simulation has no z velocity; leave uz at 0")
        return (ux, uy)

    def initialize(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        p = self.params
        rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), p.shape)
        u = [
            np.broadcast_to(np.asarray(u_a, dtype=np.float64), p.shape) + 0.5 * f_a / rho
            for u_a, f_a in zip(self._velocity_components(ux, uy, uz), p.force, strict=True)
        ]
        delta_rho = np.asarray(rho - 1.0, dtype=p.dtype)
        self._f = equilibrium(self.lat, delta_rho, tuple(np.asarray(a, dtype=p.dtype) for a in u))
        self.time_step = 0

    def _collision_force(self):
        if self._field is None:
            return self.params.force
        return tuple(f_a + self._field[a] for a, f_a in enumerate(self.params.force))

    def step(self, steps: int = 1):
        lat = self.lat
        omega_plus, omega_minus = self._rates
        for _ in range(steps):
            force = self.params.force
            if self._ibm is not None:
                local = self._ibm.force_field(self._f.reshape(-1))
                self._field.reshape(lat.d, -1)[:, self._ibm.nodes] = local
                force = self._collision_force()
            streamed = pull_stream(lat, self._f)
            streamed = apply_bounce_back(lat, streamed, self._f, self._mask, self._wall_term)
            if self._rules is not None:
                target, slots, coeff, constant = self._rules
                streamed.reshape(-1)[target] = evaluate_affine(slots, coeff, constant, self._f.reshape(-1))
            post = collide(lat, streamed, omega_plus, omega_minus, force)
            self._f = np.where(self._fluid, post, self._f)

    _advance = step

    def run(self, steps, *, every=0, callback=None, check_every=500):
        if steps < 0 or every < 0 or check_every <= 0:
            raise ValueError(f"need steps >= 0, every >= 0, check_every > 0; got {steps}, {every}, {check_every}")
        done = 0
        while done < steps:
            stop = min(steps, (done // check_every + 1) * check_every)
            if every:
                stop = min(stop, (done // every + 1) * every)
            self.step(stop - done)
            self.time_step += stop - done
            done = stop
            self._check_health()
            if callback is not None and ((every and done % every == 0) or done == steps):
                callback(self)

    def _check_health(self):
        delta_rho = self._f.sum(axis=0)
        fluid = self._fluid
        non_finite = ~np.isfinite(delta_rho) & fluid
        bad = non_finite | ((delta_rho <= -1.0) & fluid)
        if not bool(bad.any()):
            return
        location = tuple(int(i) for i in np.unravel_index(int(np.argmax(bad)), self.params.shape))
        reason = "non-finite populations" if bool(non_finite.any()) else "non-positive density"
        raise SimulationDiverged(self.time_step, location, reason)

    def populations(self):
        return np.array(self._f, copy=True)

    def evaluate_affine(self, compiled):
        dtype = self.params.dtype
        values = evaluate_affine(
            np.asarray(compiled.slots),
            np.asarray(compiled.coeff, dtype=dtype),
            np.asarray(compiled.constant, dtype=dtype),
            self._f.reshape(-1),
        )
        return values.astype(np.float64)

    @property
    def immersed_force(self):
        if self._ibm is None:
            raise AttributeError("this simulation has no immersed body")
        return np.asarray(self._ibm.body_force).astype(np.float64)

    def macroscopic(self) -> Fields:
        p = self.params
        rho, u = macroscopic_from_post_collision(self.lat, self._f, self._collision_force(), p.solid)
        return Fields(rho, u, np.array(p.solid))
