"""Float64 torch re-expression of the oracle step, differentiable in the controls.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of the HIP adjoint).

The same step as ``oracle.reference_solver.ReferenceSolver.step`` (which restates
src/lbm/solver.py:257-274 of the reference), written with torch operations on CPU float64
so that torch.autograd differentiates it:

    IBM force h(f; X) and body force from the stored state  (ibm.py:189-214)
    pull stream + bounce-back + link rules                  (streaming.py, boundaries.py, link_rules.py)
    TRT + Guo collision with F = F0 + g + h at fluid nodes  (collision.py, forcing.py)

The controls are a dense body-force field g (d, *shape) and the marker positions X
(n, d); the Roma weights are recomputed from X with the stencil base frozen at the
initial markers. The loss is the mean x body force over the last ``window`` steps.
"""

from __future__ import annotations

import numpy as np
import torch

from oracle.reference_solver import Lattice, bounce_back_links, dot, neighbour


def roma_t(r):
    a = r.abs()
    inner = (1.0 + torch.sqrt(torch.clamp(1.0 - 3.0 * a * a, min=1e-30))) / 3.0
    outer = (5.0 - 3.0 * a - torch.sqrt(torch.clamp(1.0 - 3.0 * (1.0 - a) ** 2, min=1e-30))) / 6.0
    zero = torch.zeros_like(a)
    return torch.where(a <= 0.5, inner, torch.where(a <= 1.5, outer, zero))


def equilibrium_t(lat: Lattice, delta_rho, u):
    """oracle.reference_solver.equilibrium with torch.stack (same expression order)."""
    rho = 1.0 + delta_rho
    usq = 1.5 * sum(u_a * u_a for u_a in u)
    terms = []
    for c_i, w_i in zip(lat.c, lat.w, strict=True):
        cu = 3.0 * dot(c_i, u)
        terms.append(w_i * (delta_rho + rho * (cu + 0.5 * cu * cu - usq)))
    return torch.stack([t if torch.is_tensor(t) else torch.full_like(delta_rho, t) for t in terms])


class TorchDragModel:
    def __init__(self, params, steps: int, window: int, init_state: np.ndarray):
        self.p = params
        self.lat = lat = Lattice(params.lattice.name)
        self.steps, self.window = steps, window
        self.f0 = torch.as_tensor(np.asarray(init_state, dtype=np.float64))
        shape = params.shape
        self.shape = shape
        n = int(np.prod(shape))
        self.n = n
        mask, term = bounce_back_links(lat, params.solid, params.wall_velocity, np.float64)
        self.mask = torch.as_tensor(mask)
        self.wall_term = torch.as_tensor(term)
        self.fluid = torch.as_tensor(params.fluid)
        self.opp = list(lat.opposite)
        wp, wm = params.relaxation_rates
        self.wp, self.wm = wp, wm
        rules = params.link_rules
        self.rules = None
        if rules is not None and len(rules):
            v = rules.values
            self.rules = (
                torch.as_tensor(np.asarray(rules.direction) * n + np.asarray(rules.node)),
                torch.as_tensor(v.direction * n + v.node),
                torch.as_tensor(v.coeff),
                torch.as_tensor(v.constant),
            )
        body = params.immersed
        d = lat.d
        base = np.floor(np.asarray(body.markers) - 0.5).astype(np.int64)
        offsets = np.stack(np.meshgrid(*(np.arange(3),) * d, indexing="ij"), axis=-1).reshape(-1, d)
        points = base[:, None, :] + offsets[None, :, :]
        stencil = np.ravel_multi_index(tuple(points[..., a] for a in range(d))[::-1], shape)
        nodes, inverse = np.unique(stencil, return_inverse=True)
        self.points = torch.as_tensor(points.astype(np.float64))
        self.nodes = torch.as_tensor(nodes)
        self.inverse = torch.as_tensor(inverse.reshape(stencil.shape))
        self.areas = torch.as_tensor(np.asarray(body.areas, dtype=np.float64))
        self.iterations = body.iterations
        sources = [neighbour(nodes, [-a for a in c], shape) for c in lat.c]
        self.sources = torch.as_tensor(np.stack(sources))  # (q, nn)
        self.force0 = tuple(float(f) for f in params.force)

    def _moments(self, f_flat):
        lat, n = self.lat, self.n
        gathered = torch.stack([f_flat[k * n + self.sources[k]] for k in range(lat.q)])  # (q, nn)
        drho = gathered.sum(0)
        mom = [sum(c[a] * gathered[k] for k, c in enumerate(lat.c) if c[a]) for a in range(lat.d)]
        return drho, mom

    def ibm(self, f, markers):
        lat = self.lat
        d = lat.d
        weights = torch.prod(roma_t(self.points - markers[:, None, :]), dim=-1)  # (n_m, S)
        drho, mom = self._moments(f.reshape(-1))
        rho = 1.0 + drho
        u_star = torch.stack([(mom[a] + 0.5 * self.force0[a]) / rho for a in range(d)])
        rho_m = (weights * rho[self.inverse]).sum(1)
        force = torch.zeros_like(u_star)
        total = torch.zeros(d, dtype=torch.float64)
        flat = self.inverse.reshape(-1)
        for _ in range(self.iterations):
            u = u_star + force / (2.0 * rho)
            u_m = (weights[None] * u[:, self.inverse]).sum(2)
            f_m = 2.0 * rho_m * (0.0 - u_m) * self.areas
            total = total + f_m.sum(1)
            spread = (weights[None] * f_m[:, :, None]).reshape(d, -1)
            force = force + torch.stack([torch.zeros(len(self.nodes), dtype=torch.float64).index_add(0, flat, spread[a])
                                         for a in range(d)])
        return force, -total

    def collide(self, f, F):
        lat = self.lat
        drho = f.sum(0)
        rho = 1.0 + drho
        j = [sum(c[a] * f[i] for i, c in enumerate(lat.c) if c[a]) for a in range(lat.d)]
        u = [(j[a] + 0.5 * F[a]) / rho for a in range(lat.d)]
        feq = equilibrium_t(lat, drho, u)
        neq = f - feq
        neq_o = neq[self.opp]
        post = f - 0.5 * self.wp * (neq + neq_o) - 0.5 * self.wm * (neq - neq_o)
        a_p, a_m = 1.0 - 0.5 * self.wp, 1.0 - 0.5 * self.wm
        udf = sum(u[a] * F[a] for a in range(lat.d))
        src = []
        for c, w in zip(lat.c, lat.w, strict=True):
            cf = sum(c[a] * F[a] for a in range(lat.d) if c[a]) if any(c) else 0.0
            cu = sum(c[a] * u[a] for a in range(lat.d) if c[a]) if any(c) else 0.0
            src.append(w * (a_p * (9.0 * cu * cf - 3.0 * udf) + a_m * 3.0 * cf))
        return post + torch.stack(src)

    def loss(self, g, markers):
        lat = self.lat
        d = lat.d
        f = self.f0.clone()
        axes = tuple(range(d))
        drags = []
        for t in range(self.steps):
            h_nodes, body_force = self.ibm(f, markers)
            h = torch.zeros((d, self.n), dtype=torch.float64)
            h = h.index_copy(1, self.nodes, h_nodes)
            F = [self.force0[a] + g[a] + h[a].reshape(self.shape) for a in range(d)]
            streamed = torch.stack([torch.roll(f[i], shifts=tuple(reversed(c)), dims=axes) for i, c in enumerate(lat.c)])
            streamed = torch.where(self.mask, f[self.opp] + self.wall_term, streamed)
            if self.rules is not None:
                target, slots, coeff, const = self.rules
                vals = const + (coeff * f.reshape(-1)[slots]).sum(1)
                streamed = streamed.reshape(-1).index_copy(0, target, vals).reshape(streamed.shape)
            post = self.collide(streamed, F)
            f = torch.where(self.fluid, post, f)
            if t >= self.steps - self.window:
                drags.append(body_force[0])
        return torch.stack(drags).mean()
