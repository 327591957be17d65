"""CPU oracle for the host preprocessing: node-by-node Zou-He rules, whole-grid voxels and wall links.

TEST INFRASTRUCTURE ONLY. Only ``tests/`` may import this module, as the checker of the
``lbm`` package's vectorised builders (``lbm.open_boundaries.open_channel``,
``lbm.geometry.voxelize`` / ``wall_links``); the ``lbm`` package never imports it.

It restates the reference's builders the slow, literal way, one boundary node and one
term at a time, so that a vectorised builder that groups nodes, restricts work to a
bounding box or reorders its loops is checked term for term:

* linear expressions over streamed populations ... src/lbm/open_boundaries.py:53-79
* streamed population next to resting walls ...... src/lbm/open_boundaries.py:80-86
* boundary column, Zou-He rules per node .......... src/lbm/open_boundaries.py:88-155
* open channel (close the end layers) ............. src/lbm/open_boundaries.py:156-186
* node positions, voxelisation .................... src/lbm/geometry.py:136-146
* wall links over every fluid node ................ src/lbm/geometry.py:174-206

Rows come out as lists of (direction, node, coefficient) in the reference's insertion
order with exact-zero terms dropped, and constants for shifted storage f~ = f - w, so a
comparison can be bitwise. Geometric predicates (``contains``, ``ray_fraction``) are the
body's own; what is restated is which nodes and links they are asked about.

Parity status: pinned by tests/test_golden.py's Zou-He rule rows generated from the
reference (2D and 3D), which this module reproduces bit for bit
(tests/test_preprocessing.py), and by the exact-imposition invariant of the rules.
"""

from __future__ import annotations

import numpy as np

from oracle.reference_solver import Lattice, neighbour


class Linear:
    """const + sum of coeff * f_k(y) over unshifted populations, keyed (k, y) in insertion order."""

    def __init__(self, terms=None, const: float = 0.0):
        self.terms = dict(terms or {})
        self.const = const

    def __add__(self, other: "Linear") -> "Linear":
        out = Linear(self.terms, self.const + other.const)
        for key, a in other.terms.items():
            out.terms[key] = out.terms.get(key, 0.0) + a
        return out

    def scaled(self, factor: float) -> "Linear":
        return Linear({key: a * factor for key, a in self.terms.items()}, self.const * factor)

    def shifted(self, lat: Lattice, target: int):
        """(row, constant) of the rule for f~_target with every term in shifted storage."""
        const = self.const - lat.w[target]
        row = []
        for (k, y), a in self.terms.items():
            if a != 0.0:
                row.append((k, y, a))
                const += a * lat.w[k]
        return row, const


def streamed(lat: Lattice, k: int, x: int, solid_flat, shape) -> Linear:
    """f_k(x, t+1): pulled from x - c_k, or bounced back at x when x - c_k is solid."""
    source = int(neighbour([x], [-a for a in lat.c[k]], shape)[0])
    if solid_flat[source]:
        return Linear({(lat.opposite[k], x): 1.0})
    return Linear({(k, source): 1.0})


def boundary_column(solid: np.ndarray, column: int) -> list[int]:
    """Flat indices of the fluid nodes at x = column, in flat order."""
    shape = solid.shape
    return [int(np.ravel_multi_index((*idx, column), shape)) for idx in np.ndindex(*shape[:-1])
            if not solid[(*idx, column)]]


def zou_he_rules(lat: Lattice, solid, layer, column: int, sign: int, velocity=None, rho_out=None):
    """(nodes, directions, rows, constants) of the Zou-He rules at x = column."""
    shape = solid.shape
    solid_flat = solid.ravel()
    layer_flat = layer.ravel()
    c_n = [sign * v[0] for v in lat.c]
    parallel = [k for k in range(lat.q) if c_n[k] == 0]
    inward = [k for k in range(lat.q) if c_n[k] == -1]
    unknown = [k for k in range(lat.q) if c_n[k] == 1]
    tangents = range(1, lat.d)
    nodes, directions, rows, constants = [], [], [], []
    for x in boundary_column(solid, column):
        f = {k: streamed(lat, k, x, solid_flat, shape) for k in range(lat.q)}
        s_par = Linear()
        for k in parallel:
            s_par = s_par + f[k]
        s_in = Linear()
        for k in inward:
            s_in = s_in + f[k]
        known = s_par + s_in.scaled(2.0)
        if velocity is not None:
            u = velocity[(slice(None), *np.unravel_index(x, shape))]
            rho = known.scaled(1.0 / (1.0 - sign * u[0]))
            rho_u = [rho.scaled(float(u_a)) for u_a in u]
        else:
            rho = Linear(const=rho_out)
            rho_u = [(rho + known.scaled(-1.0)).scaled(float(sign))] + [Linear() for _ in tangents]
        transverse = {}
        for t in tangents:
            half = Linear()
            for k in parallel:
                if lat.c[k][t]:
                    half = half + f[k].scaled(0.5 * lat.c[k][t])
            transverse[t] = half + rho_u[t].scaled(-1.0 / 3.0)
        sources = [int(neighbour([x], [-a for a in lat.c[j]], shape)[0]) for j in unknown]
        if not any(layer_flat[s] for s in sources):
            continue
        for j in unknown:
            value = f[lat.opposite[j]]
            for a in range(lat.d):
                if lat.c[j][a]:
                    value = value + rho_u[a].scaled(6.0 * lat.w[j] * lat.c[j][a])
            for t in tangents:
                if lat.c[j][t]:
                    value = value + transverse[t].scaled(-float(lat.c[j][t]))
            row, const = value.shifted(lat, j)
            nodes.append(x)
            directions.append(j)
            rows.append(row)
            constants.append(const)
    return nodes, directions, rows, constants


def open_channel(lat: Lattice, solid, inflow, rho_out: float = 1.0):
    """(solid with closed end layers, nodes, directions, rows, constants): velocity inlet at
    x = 1, then pressure outlet at x = nx - 2."""
    solid = np.array(solid, dtype=bool)
    inlet = np.zeros_like(solid)
    outlet = np.zeros_like(solid)
    inlet[..., 0] = ~solid[..., 0]
    outlet[..., -1] = ~solid[..., -1]
    solid[..., 0] = True
    solid[..., -1] = True
    velocity = np.zeros((lat.d, *solid.shape))
    velocity[..., 1] = np.broadcast_to(np.asarray(inflow, dtype=np.float64), velocity[..., 1].shape)
    a = zou_he_rules(lat, solid, inlet, 1, +1, velocity=velocity)
    b = zou_he_rules(lat, solid, outlet, solid.shape[-1] - 2, -1, rho_out=rho_out)
    return (solid, *(p + q for p, q in zip(a, b, strict=True)))


def node_positions(shape) -> np.ndarray:
    """Coordinates (x, y[, z]) of every node, shape (*shape, d)."""
    axes = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    return np.stack(axes[::-1], axis=-1)


def voxelize(body, shape) -> np.ndarray:
    """Mask of the nodes inside ``body``, every node of the grid asked."""
    return body.contains(node_positions(shape))


def wall_links(lat: Lattice, body, solid):
    """(node, direction, q) of every link from a fluid node x into ``body`` (x - c_j inside),
    directions in lattice order, nodes in flat order within a direction."""
    shape = solid.shape
    inside = voxelize(body, shape).ravel()
    fluid = np.flatnonzero(~solid.ravel())
    positions = node_positions(shape).reshape(-1, len(shape))
    nodes, directions, qs = [], [], []
    for j, c_j in enumerate(lat.c):
        if not any(c_j):
            continue
        up = neighbour(fluid, [-a for a in c_j], shape)
        x = fluid[inside[up]]
        if x.size:
            step = np.broadcast_to(-np.asarray(c_j, dtype=np.float64), (x.size, len(shape)))
            nodes.append(x)
            directions.append(np.full(x.size, j, dtype=np.int64))
            qs.append(body.ray_fraction(positions[x], step))
    if not nodes:
        return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64), np.zeros(0)
    return np.concatenate(nodes), np.concatenate(directions), np.concatenate(qs)
