"""Momentum-exchange force on a body, as a fixed affine functional of the stored state.

For every wall link (fluid x, incoming direction j, body node x - c_j) with
i = opp(j) pointing into the body (the functional of the reference's
``src/lbm/forces.py:1-103``)::

    F_body = sum_links c_i [ f*_i(x, t) + f_j(x, t+1) ]

f_j(x, t+1) is what the boundary makes of f*_i: plain bounce-back
(f*_i(x) + 6 w_j c_j . u_w) or the link's rule row. With shifted storage each
link also carries the constant 2 w_i. The rows are assembled here for all links
at once (array operations over the link list); the solver evaluates them in its
own layout, on the device (``HipSolver.evaluate_affine`` / ``force_on_device``).

Row format (fixed, since the device sums a row's terms in order): the term
(i, x, 1) first; then, for a plain link, (i, x, 1) again, or the nonzero terms of
the link's rule in the rule's order; zero padding to the widest row.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from lbm.geometry import WallLinks, neighbour_index
from lbm.link_rules import AffineSet, CompiledAffine, compile_affine
from lbm.params import SimulationParams


def _rule_index(params: SimulationParams, node: np.ndarray, direction: np.ndarray) -> np.ndarray:
    """Index of the link rule on each (node, direction) link, -1 where there is none."""
    rules = params.link_rules
    found = np.full(node.size, -1, dtype=np.int64)
    if rules is None or not len(rules) or not node.size:
        return found
    q = params.lattice.q
    keys = np.asarray(rules.node, dtype=np.int64) * q + np.asarray(rules.direction, dtype=np.int64)
    order = np.argsort(keys)
    wanted = node * q + direction
    at = np.minimum(np.searchsorted(keys[order], wanted), keys.size - 1)
    hit = keys[order][at] == wanted
    found[hit] = order[at[hit]]
    return found


def _probe_rows(params: SimulationParams, links: WallLinks) -> AffineSet:
    """The affine rows f*_i(x) + f_j(x, t+1) + 2 w_i of every link, in link order."""
    lattice = params.lattice
    c = np.asarray(lattice.c, dtype=np.int64)
    w = np.asarray(lattice.w, dtype=np.float64)
    node = np.asarray(links.node, dtype=np.int64)
    j = np.asarray(links.direction, dtype=np.int64)
    i = np.asarray(lattice.opposite, dtype=np.int64)[j]
    rule = _rule_index(params, node, j)
    plain = rule < 0

    # Plain bounce-back: a second (i, x, 1) term and the moving-wall constant of the
    # wall node x - c_j, 2 w_j / cs^2 (c_j . u_w).
    constant = 2.0 * w[i]
    if plain.any():
        source = neighbour_index(node[plain], -c[j[plain]].T, params.shape)
        walls = params.wall_velocity.reshape(lattice.d, -1)[:, source]
        c_dot_u = np.einsum("ld,dl->l", c[j[plain]].astype(np.float64), walls)
        constant[plain] += 2.0 * w[j[plain]] / lattice.cs2 * c_dot_u

    # Rule links: the rule's nonzero terms after (i, x, 1), in the rule's order.
    ruled = np.flatnonzero(~plain)
    if ruled.size:
        values = params.link_rules.values
        coeff = values.coeff[rule[ruled]]
        keep = coeff != 0.0
        counts = keep.sum(axis=1)
        constant[ruled] += values.constant[rule[ruled]]
    width = max([2] * bool(plain.any()) + ([1 + int(counts.max())] if ruled.size else [1]))

    n = node.size
    direction = np.zeros((n, width), dtype=np.int64)
    nodes = np.zeros((n, width), dtype=np.int64)
    coefficient = np.zeros((n, width))
    direction[:, 0], nodes[:, 0], coefficient[:, 0] = i, node, 1.0
    direction[plain, 1], nodes[plain, 1], coefficient[plain, 1] = i[plain], node[plain], 1.0
    if ruled.size:
        row, term = np.nonzero(keep)  # row-major: each row's kept terms in order
        column = 1 + np.arange(row.size) - np.repeat(np.cumsum(counts) - counts, counts)
        target = ruled[row]
        source_rows = rule[ruled][row]
        direction[target, column] = values.direction[source_rows, term]
        nodes[target, column] = values.node[source_rows, term]
        coefficient[target, column] = values.coeff[source_rows, term]
    return AffineSet(direction, nodes, coefficient, constant)


@dataclass(frozen=True, eq=False)
class ForceProbe:
    """Force = weight.T @ values: weight[r] = -c_j (= c_i) of link r, values the rows above."""

    params: SimulationParams
    weight: np.ndarray
    values: AffineSet
    _compiled: dict = field(default_factory=dict, repr=False)

    @staticmethod
    def build(params: SimulationParams, links: WallLinks) -> ForceProbe:
        weight = -np.asarray(params.lattice.c, dtype=np.float64)[np.asarray(links.direction, dtype=np.int64)]
        return ForceProbe(params=params, weight=weight.reshape(-1, params.d), values=_probe_rows(params, links))

    def compiled(self, layout: str) -> CompiledAffine:
        """The rows resolved to slots of ``layout`` (built once per layout)."""
        found = self._compiled.get(layout)
        if found is None:
            p = self.params
            found = self._compiled[layout] = compile_affine(p.lattice, self.values, layout, p.solid, p.wall_velocity)
        return found

    def evaluate(self, solver) -> np.ndarray:
        """Force on the body from the solver's current state, shape (d,), host float64."""
        if solver.params is not self.params:
            raise ValueError("this probe was built for a different SimulationParams")
        return self.weight.T @ solver.evaluate_affine(self.compiled(solver.state_layout))
