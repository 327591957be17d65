"""Momentum-exchange force on a body, as a fixed affine functional of the stored state.

For every wall link (fluid x, incoming direction j, body node x - c_j) with
i = opp(j) pointing into the body (``src/lbm/forces.py:1-103`` of the reference)::

    F_body = sum_links c_i [ f*_i(x, t) + f_j(x, t+1) ]

f_j(x, t+1) is what the boundary makes of f*_i: plain bounce-back
(f*_i(x) + 6 w_j c_j . u_w) or the link's rule row. With shifted storage each
link also carries the constant 2 w_i. The functional is built once per body;
the solver evaluates it on the device in its own layout
(``HipSolver.evaluate_affine`` / ``HipSolver.force``), one sparse gather per call.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from lbm.geometry import WallLinks, neighbour_index
from lbm.link_rules import AffineSet, CompiledAffine, compile_affine
from lbm.params import SimulationParams


@dataclass(frozen=True, eq=False)
class ForceProbe:
    """Force = weight.T @ values, weight[r] = -c_j = c_i of link r, values an ``AffineSet``."""

    params: SimulationParams
    weight: np.ndarray
    values: AffineSet
    _compiled: dict = field(default_factory=dict, repr=False)

    @staticmethod
    def build(params: SimulationParams, links: WallLinks) -> ForceProbe:
        lattice = params.lattice
        shape = params.shape
        opposite = np.asarray(lattice.opposite)
        rules = params.link_rules
        rule_of = {}
        if rules is not None:
            rule_of = {key: r for r, key in enumerate(zip(rules.node.tolist(), rules.direction.tolist(), strict=True))}
        walls = params.wall_velocity.reshape(lattice.d, -1)
        rows, constants = [], []
        for x, j in zip(links.node.tolist(), links.direction.tolist(), strict=True):
            i = int(opposite[j])
            row = [(i, x, 1.0)]
            constant = 2.0 * lattice.w[i]
            r = rule_of.get((x, j))
            if r is None:
                source = int(neighbour_index([x], [-a for a in lattice.c[j]], shape)[0])
                row.append((i, x, 1.0))
                constant += 2.0 * lattice.w[j] / lattice.cs2 * float(np.dot(lattice.c[j], walls[:, source]))
            else:
                v = rules.values
                row += [
                    (int(k), int(y), float(a))
                    for k, y, a in zip(v.direction[r], v.node[r], v.coeff[r], strict=True)
                    if a != 0.0
                ]
                constant += float(v.constant[r])
            rows.append(row)
            constants.append(constant)
        weight = -np.asarray([lattice.c[j] for j in links.direction], dtype=np.float64)
        return ForceProbe(params=params, weight=weight.reshape(-1, lattice.d), values=AffineSet.from_lists(rows, constants))

    def compiled(self, layout: str) -> CompiledAffine:
        if layout not in self._compiled:
            p = self.params
            self._compiled[layout] = compile_affine(p.lattice, self.values, layout, p.solid, p.wall_velocity)
        return self._compiled[layout]

    def evaluate(self, solver) -> np.ndarray:
        """Force on the body from the solver's current state, shape (d,), host float64."""
        if solver.params is not self.params:
            raise ValueError("this probe was built for a different SimulationParams")
        values = solver.evaluate_affine(self.compiled(solver.state_layout))
        return self.weight.T @ values
