"""Affine link rules and their compilation to flat slots of each memory layout.

A rule sets one streamed population from post-collision populations of the
previous step (``src/lbm/link_rules.py:1-40`` of the reference)::

    f_j(x, t+1) = b + sum_m a_m f*_{k_m}(y_m, t)

Rules are described geometrically (direction, node) and *compiled* per layout
into flat slots of the (q, N) population buffer, N = number of nodes:

=============  ===============================================  ==================
layout         where f*_k(y) is stored                          slot a rule writes
=============  ===============================================  ==================
``reference``  (k, y)                                           (j, x)
``pull``       (k, y)                                           (opp j, x)
``aa_odd``     (opp k, y)                                       (j, x)
``aa_even``    (k, y + c_k); (opp k, y) if y + c_k is solid     (j, x)
=============  ===============================================  ==================

The written slot is the one the stream-collide kernel's bounce-back branch
reads, so kernels need no rule awareness. The device pre-pass evaluates every
rule before writing any (see ``csrc/lbm_rules.hip``), or, for a set in which no
rule reads a slot another rule writes (``reads_other_targets``), evaluates and
writes each rule in one launch with the same result.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from lbm.geometry import WallLinks, neighbour_index
from lbm.lattice import VelocitySet

LAYOUTS = ("reference", "pull", "aa_odd", "aa_even")


def _check_layout(layout: str) -> None:
    if layout not in LAYOUTS:
        raise ValueError(f"unknown layout {layout!r}; choose from {LAYOUTS}")


@dataclass(frozen=True)
class AffineSet:
    """n affine expressions, each padded to m terms (``link_rules.py:54-103``).

    value[r] = constant[r] + sum_t coeff[r, t] * f*_{direction[r, t]}(node[r, t]);
    padding terms have coefficient 0 and point at direction 0 of node 0.
    """

    direction: np.ndarray
    node: np.ndarray
    coeff: np.ndarray
    constant: np.ndarray

    def __len__(self) -> int:
        return len(self.constant)

    @property
    def terms(self) -> int:
        return self.coeff.shape[1]

    @staticmethod
    def from_lists(rows, constant=None) -> AffineSet:
        """Build from rows of (direction, node, coefficient) triples."""
        n = len(rows)
        m = max((len(row) for row in rows), default=1)
        direction = np.zeros((n, m), dtype=np.int64)
        node = np.zeros((n, m), dtype=np.int64)
        coeff = np.zeros((n, m))
        for r, row in enumerate(rows):
            if row:
                k, y, a = zip(*row, strict=True)
                direction[r, : len(row)] = k
                node[r, : len(row)] = y
                coeff[r, : len(row)] = a
        const = np.zeros(n) if constant is None else np.asarray(constant, dtype=np.float64)
        return AffineSet(direction, node, coeff, const)

    @staticmethod
    def concat(*sets: AffineSet) -> AffineSet:
        sets = tuple(s for s in sets if len(s))
        if not sets:
            return AffineSet.from_lists([])
        m = max(s.terms for s in sets)

        def widen(a):
            return np.pad(a, ((0, 0), (0, m - a.shape[1])))

        return AffineSet(
            np.concatenate([widen(s.direction) for s in sets]),
            np.concatenate([widen(s.node) for s in sets]),
            np.concatenate([widen(s.coeff) for s in sets]),
            np.concatenate([s.constant for s in sets]),
        )


@dataclass(frozen=True)
class LinkRules:
    """Rules f_j(x, t+1) = values[r] on links (node[r], direction[r]); one rule per link."""

    node: np.ndarray
    direction: np.ndarray
    values: AffineSet

    def __post_init__(self) -> None:
        if not (len(self.node) == len(self.direction) == len(self.values)):
            raise ValueError("node, direction and values must have the same length")
        if len(self.node):
            node = np.asarray(self.node, dtype=np.int64)
            direction = np.asarray(self.direction, dtype=np.int64)
            if direction.min() < 0 or node.min() < 0:
                raise ValueError("link nodes and directions must be non-negative")
            key = np.sort(node * (int(direction.max()) + 1) + direction)  # one int64 per link
            if np.any(key[1:] == key[:-1]):
                raise ValueError("a link has more than one rule")

    def __len__(self) -> int:
        return len(self.node)

    @staticmethod
    def concat(*rules: LinkRules) -> LinkRules:
        rules = tuple(r for r in rules if r is not None)
        return LinkRules(
            np.concatenate([r.node for r in rules]).astype(np.int64),
            np.concatenate([r.direction for r in rules]).astype(np.int64),
            AffineSet.concat(*(r.values for r in rules)),
        )

    def validate(self, lattice: VelocitySet, solid: np.ndarray, wall_velocity: np.ndarray) -> None:
        """Every rule must sit on a link from a fluid node to a solid node at rest."""
        shape = solid.shape
        flat_solid = solid.ravel()
        if np.any(self.direction >= lattice.q) or np.any(self.direction < 0):
            raise ValueError(f"rule directions must be in [0, {lattice.q})")
        if np.any(flat_solid[self.node]):
            raise ValueError("link rules must target fluid nodes")
        walls = wall_velocity.reshape(lattice.d, -1)
        for j in np.unique(self.direction):
            upstream = neighbour_index(self.node[self.direction == j], [-a for a in lattice.c[j]], shape)
            if not np.all(flat_solid[upstream]):
                raise ValueError(f"a rule for direction {j} does not start at a solid node")
            if np.abs(walls[:, upstream]).max(initial=0.0) > 0:
                raise ValueError("link rules are only allowed next to solid nodes at rest")


def bouzidi_rules(lattice: VelocitySet, links: WallLinks, solid: np.ndarray) -> LinkRules:
    """Linear interpolated bounce-back (Bouzidi, Firdaouss & Lallemand 2001).

    For the link from fluid x towards body node x - c_j with wall fraction q
    and i = opp(j) (``link_rules.py:154-188``)::

        q <  1/2:  f_j(x) = 2q f*_i(x) + (1 - 2q) f*_i(x + c_j)
        q >= 1/2:  f_j(x) = f*_i(x) / (2q) + (1 - 1/(2q)) f*_j(x)

    A q < 1/2 link whose x + c_j is solid keeps plain bounce-back (no rule).
    """
    shape = solid.shape
    flat_solid = np.asarray(solid, dtype=bool).ravel()
    x = np.asarray(links.node, dtype=np.int64)
    j = np.asarray(links.direction, dtype=np.int64)
    q = np.asarray(links.q, dtype=np.float64)
    i = np.asarray(lattice.opposite)[j]
    behind = np.empty_like(x)
    for jj in np.unique(j):
        sel = j == jj
        behind[sel] = neighbour_index(x[sel], lattice.c[jj], shape)
    near = q < 0.5
    keep = ~near | ~flat_solid[behind]
    x, j, q, i, behind, near = x[keep], j[keep], q[keep], i[keep], behind[keep], near[keep]
    n = len(x)
    direction = np.empty((n, 2), dtype=np.int64)
    node = np.empty((n, 2), dtype=np.int64)
    coeff = np.empty((n, 2))
    direction[:, 0] = i
    node[:, 0] = x
    direction[:, 1] = np.where(near, i, j)
    node[:, 1] = np.where(near, behind, x)
    with np.errstate(divide="ignore"):
        coeff[:, 0] = np.where(near, 2.0 * q, 0.5 / q)
        coeff[:, 1] = np.where(near, 1.0 - 2.0 * q, 1.0 - 0.5 / q)
    if n == 0:
        return LinkRules(x, j, AffineSet.from_lists([]))
    return LinkRules(x, j, AffineSet(direction, node, coeff, np.zeros(n)))


@dataclass(frozen=True)
class CompiledAffine:
    """An ``AffineSet`` resolved to flat slots of one layout (host arrays)."""

    slots: np.ndarray
    coeff: np.ndarray
    constant: np.ndarray


def compile_affine(
    lattice: VelocitySet, values: AffineSet, layout: str, solid: np.ndarray, wall_velocity: np.ndarray
) -> CompiledAffine:
    """Flat slots for every term of ``values`` in ``layout`` (``link_rules.py:203-240``).

    In ``aa_even`` a post-collision f*_k(y) whose push target y + c_k is
    solid was reflected into (opp k, y) together with the moving-wall term of
    direction opp k; reading it back adds the wall term of direction k to the
    constant (6 w_k c_k . u_w at the solid node).
    """
    _check_layout(layout)
    shape = solid.shape
    n = int(np.prod(shape))
    k = values.direction
    y = values.node
    constant = values.constant.copy()
    opposite = np.asarray(lattice.opposite)
    if layout in ("reference", "pull"):
        return CompiledAffine(k * n + y, values.coeff, constant)
    if layout == "aa_odd":
        return CompiledAffine(opposite[k] * n + y, values.coeff, constant)
    slots = np.empty_like(k)
    wall_terms = np.zeros(k.shape)
    flat_solid = np.asarray(solid, dtype=bool).ravel()
    walls = np.asarray(wall_velocity, dtype=np.float64).reshape(lattice.d, -1)
    for kk in np.unique(k):
        sel = k == kk
        dest = neighbour_index(y[sel], lattice.c[kk], shape)
        blocked = flat_solid[dest]
        slots[sel] = np.where(blocked, opposite[kk] * n + y[sel], kk * n + dest)
        c_dot_uw = np.tensordot(np.asarray(lattice.c[kk], dtype=np.float64), walls[:, dest], 1)
        wall_terms[sel] = (2.0 * lattice.w[kk] / lattice.cs2) * c_dot_uw * blocked
    constant += (values.coeff * wall_terms).sum(axis=1)
    return CompiledAffine(slots, values.coeff, constant)


def compile_targets(lattice: VelocitySet, rules: LinkRules, layout: str, n_nodes: int) -> np.ndarray:
    """Flat slot each rule writes (``link_rules.py:243-250``)."""
    _check_layout(layout)
    direction = np.asarray(rules.direction)
    if layout == "pull":
        direction = np.asarray(lattice.opposite)[direction]
    return direction * n_nodes + np.asarray(rules.node)


@dataclass(frozen=True)
class CompiledRules:
    target: np.ndarray
    values: CompiledAffine


def compile_rules(lattice: VelocitySet, rules: LinkRules, layout: str, solid, wall_velocity) -> CompiledRules:
    n = int(np.prod(solid.shape))
    return CompiledRules(
        target=compile_targets(lattice, rules, layout, n),
        values=compile_affine(lattice, rules.values, layout, solid, wall_velocity),
    )


def reads_other_targets(target, slots, coeff) -> bool:
    """Whether some rule has a nonzero term on a slot that a *different* rule writes.

    When it does not, evaluating and writing each rule in one pass, in any order, gives the
    values of the two-phase pre-pass (every rule reads pre-rule populations; a rule that
    reads its own target does so before it writes it), so the device can apply the set in
    one launch. ``slots``/``coeff`` are (n_rules, m); zero coefficients are padding.
    """
    target = np.asarray(target, dtype=np.int64)
    if target.size == 0:
        return False
    coeff = np.asarray(coeff)
    rows, terms = np.nonzero(coeff != 0)
    read = np.asarray(slots, dtype=np.int64)[rows, terms]
    order = np.argsort(target, kind="stable")
    written = target[order]
    at = np.minimum(np.searchsorted(written, read), written.size - 1)
    hit = written[at] == read
    return bool(np.any(order[at[hit]] != rows[hit]))
