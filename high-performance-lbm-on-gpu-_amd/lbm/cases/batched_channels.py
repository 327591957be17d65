"""A seeded family of independent channel flows around Bouzidi cylinders (the batched workload).

Each channel: height H (walls included, H in [64, 512]), length 4H, a cylinder of radius
R in [4, 24] (at most H/5) with interpolated (Bouzidi) bounce-back at a seeded off-axis
position, a parabolic Zou-He velocity inlet (u_max = 0.075) and pressure outlet, and
Re = u_mean 2R / nu in [20, 200] setting tau. D2Q9, float32 by default.
"""

from __future__ import annotations

import numpy as np

from lbm.boundaries import channel_walls
from lbm.cases.ibm_channel import inflow_profile
from lbm.geometry import circle, voxelize, wall_links
from lbm.lattice import CS2, D2Q9
from lbm.link_rules import LinkRules, bouzidi_rules
from lbm.open_boundaries import open_channel
from lbm.params import SimulationParams

U_MAX = 0.075


def channel_with_cylinder(height: int, radius: float, center, reynolds: float, dtype="float32"):
    nx, ny = 4 * height, height
    solid, open_rules = open_channel(D2Q9, channel_walls(nx, ny), inflow_profile(ny, U_MAX))
    body = circle(*center, radius)
    solid = solid | voxelize(body, solid.shape)
    links = wall_links(D2Q9, body, solid)
    rules = LinkRules.concat(open_rules, bouzidi_rules(D2Q9, links, solid))
    nu = (2.0 / 3.0 * U_MAX) * 2.0 * radius / reynolds
    return SimulationParams(nx=nx, ny=ny, tau=nu / CS2 + 0.5, collision="trt", solid=solid, dtype=dtype,
                            link_rules=rules)


def batched_channels(count: int = 64, seed: int = 0, dtype: str = "float32", h_range=(64, 512)):
    """``count`` channel ``SimulationParams`` drawn from ``seed``."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        h = int(rng.integers(h_range[0], h_range[1] + 1))
        r = float(rng.uniform(4.0, min(24.0, h / 5.0)))
        cx = float(rng.uniform(0.5, 1.0) * h)
        cy = float(h / 2.0 + rng.uniform(-0.15, 0.15) * h)
        re = float(rng.uniform(20.0, 200.0))
        out.append(channel_with_cylinder(h, r, (cx, cy), re, dtype))
    return out


def initial_fields(params):
    profile = inflow_profile(params.ny, U_MAX)
    ux = np.broadcast_to(profile[0][:, None], params.shape).copy()
    ux[params.solid] = 0.0
    return 1.0, ux, np.zeros(params.shape)
