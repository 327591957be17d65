"""Channel with an immersed-boundary cylinder: the differentiable drag configuration.

D2Q9 channel nx x ny with no-slip walls at top and bottom, a parabolic Zou-He velocity
inlet (u_max) and a pressure outlet (rho = 1), and a cylinder of radius R represented by
IBM markers spaced ``spacing`` lattice units with ``iterations`` multi-direct passes. The
Reynolds number Re = u_mean (2R) / nu with u_mean = 2/3 u_max sets tau. Used by
``lbm.adjoint.DragAdjoint`` for gradients of the mean drag.
"""

from __future__ import annotations

import numpy as np

from lbm.boundaries import channel_walls
from lbm.geometry import circle
from lbm.ibm import ImmersedBoundary
from lbm.lattice import D2Q9, CS2
from lbm.open_boundaries import open_channel
from lbm.params import SimulationParams


def inflow_profile(ny: int, u_max: float) -> np.ndarray:
    """Parabola between the walls at y = 0.5 and ny - 1.5, shape (2, ny)."""
    y = np.arange(ny) - 0.5
    h = ny - 2
    profile = np.zeros((2, ny))
    profile[0] = np.where((y > 0) & (y < h), 4.0 * u_max * y * (h - y) / h**2, 0.0)
    return profile


def ibm_channel_params(nx=512, ny=128, radius=12.0, center=None, reynolds=100.0, u_max=0.075, spacing=1.0,
                       iterations=3, dtype="float32") -> SimulationParams:
    center = (nx / 4.0, ny / 2.0 + 0.5) if center is None else center
    u_mean = 2.0 / 3.0 * u_max
    nu = u_mean * 2.0 * radius / reynolds
    tau = nu / CS2 + 0.5
    solid, rules = open_channel(D2Q9, channel_walls(nx, ny), inflow_profile(ny, u_max))
    body = ImmersedBoundary.from_ball(circle(*center, radius), spacing=spacing, iterations=iterations)
    return SimulationParams(nx=nx, ny=ny, tau=tau, collision="trt", solid=solid, dtype=dtype, immersed=body,
                            link_rules=rules)


def initial_fields(params, u_max=0.075):
    """rho = 1 and the inlet parabola everywhere."""
    profile = inflow_profile(params.ny, u_max)
    ux = np.broadcast_to(profile[0][:, None], params.shape).copy()
    ux[params.solid] = 0.0
    return 1.0, ux, np.zeros(params.shape)
