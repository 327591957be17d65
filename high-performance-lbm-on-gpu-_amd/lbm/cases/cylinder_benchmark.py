"""Flow past a cylinder in a channel: the Schaefer-Turek DFG benchmark 2D-1 / 2D-2, and the
fixed-size drag channel (``src/lbm/cases/cylinder_benchmark.py`` of the reference).

Benchmark geometry (Schaefer & Turek 1996): channel H = 0.41 m x L = 2.2 m, cylinder of
diameter D = 0.1 m centred at (0.2, 0.2), nu = 1e-3 m^2/s, parabolic inflow
u(0, y) = 4 U_m y (H - y) / H^2 with U_mean = 2 U_m / 3 and Re = U_mean D / nu;
2D-1: U_m = 0.3 (Re 20, steady), 2D-2: U_m = 1.5 (Re 100, shedding). Reported:
C_D = 2 F_x / (rho U_mean^2 D), C_L likewise with F_y, dp = p(0.15, 0.2) - p(0.25, 0.2)
and St = D f / U_mean, with the benchmark's acceptance intervals in ``INTERVALS``.

Lattice: N nodes per diameter; Zou-He velocity inlet acting at x = 1 and pressure outlet
at x = nx - 2 (``lbm.open_boundaries.open_channel``), halfway bounce-back walls, and the
cylinder as Bouzidi interpolated bounce-back, staircase bounce-back or IBM markers.

MI355X specifics: the observer reads the momentum-exchange force into a float64 device
tensor (``lbm_affine_force``) and the two pressure points with one sparse gather, and
``force_history`` records drag and lift after every step without any host round trip
(the 880 x 164 drag channel, ``drag_channel``, records them every step for 50,000 steps).
"""

from __future__ import annotations

import argparse
import math
import time
from dataclasses import dataclass
from typing import NamedTuple

import numpy as np

from lbm.backends import make_solver
from lbm.boundaries import channel_walls
from lbm.cases.common import add_backend_arguments, even, run_until_steady
from lbm.forces import ForceProbe
from lbm.geometry import circle, voxelize, wall_links
from lbm.ibm import ImmersedBoundary
from lbm.lattice import CS2, D2Q9
from lbm.link_rules import AffineSet, LinkRules, bouzidi_rules, compile_affine
from lbm.open_boundaries import open_channel
from lbm.params import SimulationParams

HEIGHT, LENGTH, DIAMETER, CENTRE = 0.41, 2.2, 0.1, (0.2, 0.2)
VISCOSITY = 1e-3
TESTS = {"2D-1": 0.3, "2D-2": 1.5}
"""Peak inflow U_m [m/s] of each test."""
DEFAULT_U_MAX = {"2D-1": 0.01875, "2D-2": 0.05}
"""Peak inflow in lattice units: the drag error of this weakly compressible scheme grows
as Ma^2 (reference measurements at N = 40, 2D-1: C_D 5.694 / 5.609 / 5.589 for
u_max = 0.075 / 0.0375 / 0.01875)."""
INTERVALS = {
    "2D-1": {"cd": (5.57, 5.59), "cl": (0.0104, 0.0110), "dp": (0.1172, 0.1176)},
    "2D-2": {"cd": (3.22, 3.24), "cl": (0.99, 1.01), "st": (0.295, 0.305)},
}
METHODS = ("bouzidi", "staircase", "ibm")


@dataclass(frozen=True)
class Geometry:
    """The benchmark at N nodes per diameter."""

    n: int
    params: SimulationParams
    cylinder_links: object
    u_mean: float
    """Mean inflow velocity, lattice units."""
    velocity_scale: float
    """Physical velocity per lattice velocity."""
    inflow: np.ndarray
    """Inflow u_x(y), shape (ny,)."""

    def start(self, solver) -> None:
        """The inflow profile everywhere (zero inside the cylinder): no impulsive start, whose
        pressure waves would bounce between the two Zou-He ends for ~100 convective times."""
        ux = np.where(self.params.solid, 0.0, self.inflow[:, None])
        solver.initialize(1.0, ux, 0.0)


def build(test: str, n: int, u_max: float | None = None, collision: str = "trt", dtype: str = "float64",
          method: str = "bouzidi") -> Geometry:
    """Lattice, boundaries and parameters of one test at N = D / dx (``cylinder_benchmark.py:93-147``)."""
    if test not in TESTS:
        raise ValueError(f"test must be one of {tuple(TESTS)}; got {test!r}")
    if method not in METHODS:
        raise ValueError(f"method must be one of {METHODS}; got {method!r}")
    if n % 10:
        raise ValueError("N must be a multiple of 10 so that H = 4.1 D spans whole nodes")
    u_max = DEFAULT_U_MAX[test] if u_max is None else u_max
    dx = DIAMETER / n
    ny = round(HEIGHT / dx) + 2  # fluid rows 1 .. ny - 2; walls halfway at y = 0.5, ny - 1.5
    nx = round(LENGTH / dx) + 3  # inlet acts at x = 1, outlet at x = nx - 2
    height = ny - 2
    y = np.arange(ny) - 0.5
    profile = np.where((y > 0) & (y < height), 4.0 * u_max * y * (height - y) / height**2, 0.0)
    body = circle(1.0 + CENTRE[0] / dx, 0.5 + CENTRE[1] / dx, 0.5 * n)
    immersed = ImmersedBoundary.from_ball(body) if method == "ibm" else None
    solid = channel_walls(nx, ny)
    if immersed is None:
        solid |= voxelize(body, solid.shape)
    solid, rules = open_channel(D2Q9, solid, np.stack([profile, np.zeros(ny)]))
    links = wall_links(D2Q9, body, solid) if immersed is None else None
    if method == "bouzidi":
        rules = LinkRules.concat(rules, bouzidi_rules(D2Q9, links, solid))
    u_mean = 2.0 * u_max / 3.0
    reynolds = TESTS[test] * 2.0 / 3.0 * DIAMETER / VISCOSITY
    nu = u_mean * n / reynolds
    params = SimulationParams(nx=nx, ny=ny, tau=nu / CS2 + 0.5, collision=collision, solid=solid, dtype=dtype,
                              link_rules=rules, immersed=immersed)
    return Geometry(n, params, links, u_mean, TESTS[test] / u_max, profile)


def drag_channel(reynolds: float = 20.0, nx: int = 880, ny: int = 164, radius: float = 20.0,
                 centre=(200.0, 80.0), u_max: float = 0.075, collision: str = "trt", dtype: str = "float32"):
    """The fixed-size drag channel: Bouzidi cylinder of ``radius`` at ``centre`` in an
    ``nx`` x ``ny`` channel with a parabolic Zou-He inlet (peak ``u_max``), a pressure outlet
    (rho = 1) and no-slip walls; Re = U_mean 2R / nu. Returns (params, links, inflow profile)."""
    height = ny - 2
    y = np.arange(ny) - 0.5
    profile = np.where((y > 0) & (y < height), 4.0 * u_max * y * (height - y) / height**2, 0.0)
    body = circle(centre[0], centre[1], radius)
    solid = channel_walls(nx, ny) | voxelize(body, (ny, nx))
    solid, open_rules = open_channel(D2Q9, solid, np.stack([profile, np.zeros(ny)]))
    links = wall_links(D2Q9, body, solid)
    rules = LinkRules.concat(open_rules, bouzidi_rules(D2Q9, links, solid))
    nu = (2.0 * u_max / 3.0) * 2.0 * radius / reynolds
    params = SimulationParams(nx=nx, ny=ny, tau=nu / CS2 + 0.5, collision=collision, solid=solid, dtype=dtype,
                              link_rules=rules)
    return params, links, profile


def force_history(solver, probe: ForceProbe, steps: int) -> np.ndarray:
    """Advance ``steps`` steps and return the momentum-exchange force after each
    ``solver.step_multiple`` steps, shape (steps // step_multiple, d), float64.

    Every value is reduced on the device into one float64 tensor; the host sees it once at
    the end (one health check, as ``Solver.run`` with ``check_every = steps``)."""
    m = solver.step_multiple
    if steps <= 0 or steps % m:
        raise ValueError(f"steps must be a positive multiple of {m}")
    if hasattr(solver, "record_forces"):  # the same launches, prepared once
        return solver.record_forces(probe, steps)
    torch = solver._torch
    out = torch.zeros((steps // m, solver.params.d), dtype=torch.float64, device=solver.device)
    for k in range(steps // m):
        solver._advance(m)
        solver.time_step += m
        solver.force_on_device(probe, out[k])
    solver._check_health()
    return out.cpu().numpy()


@dataclass(frozen=True)
class ChannelRun:
    """A drag-channel run with the force recorded after every step (pair of AA steps)."""

    reynolds: float
    steps: int
    seconds: float
    mlups: float
    cd: float
    """Mean C_D = 2 F_x / (U_mean^2 2R) over the second half of the history."""
    cl: float
    history: np.ndarray


def run_drag_channel(reynolds: float = 20.0, steps: int = 50_000, backend: str = "hip-pull",
                     dtype: str = "float32", **geometry) -> ChannelRun:
    """The 880 x 164 Bouzidi drag channel for ``steps`` steps from the inflow profile, with
    drag and lift reduced on the device after every step; the throughput counts every node
    and includes the rule pre-pass and the per-step force reduction."""
    params, links, profile = drag_channel(reynolds=reynolds, dtype=dtype, **geometry)
    probe = ForceProbe.build(params, links)
    solver = make_solver(params, backend=backend)
    solver.initialize(1.0, np.where(params.solid, 0.0, profile[:, None]), 0.0)
    solver.synchronize()
    start = time.perf_counter()
    history = force_history(solver, probe, steps)
    seconds = time.perf_counter() - start
    radius = geometry.get("radius", 20.0)
    u_mean = 2.0 * geometry.get("u_max", 0.075) / 3.0
    coeff = history[len(history) // 2 :] / (0.5 * u_mean**2 * 2.0 * radius)
    return ChannelRun(reynolds, steps, seconds, params.n_nodes * steps / seconds / 1e6, float(coeff[:, 0].mean()),
                      float(coeff[:, 1].mean()), history)


def _pressure_points(g: Geometry) -> AffineSet:
    """Two affine rows: the mean of 1 + rho - 1 = sum_k f~*_k over the two lattice rows around
    y = 0.2, at x = 0.15 and x = 0.25 (weight 1/2 per node and direction)."""
    dx = DIAMETER / g.n
    centre_row = 0.5 + CENTRE[1] / dx
    ys = np.array([math.floor(centre_row), math.ceil(centre_row)])
    xs = np.array([round(1 + 0.15 / dx), round(1 + 0.25 / dx)])
    q = g.params.lattice.q
    nodes = np.ravel_multi_index((ys[None, :], xs[:, None]), g.params.shape)  # (point, row)
    per_point = np.repeat(nodes, q, axis=1)  # node of each term: rows outer, directions inner
    directions = np.tile(np.arange(q), (2, len(ys)))
    return AffineSet(directions, per_point.astype(np.int64), np.full(per_point.shape, 0.5), np.zeros(2))


class Observer:
    """Callable recording (C_D, C_L, dp) of a running benchmark; ``history`` holds
    (time step, C_D, C_L, dp) per call. The drag and lift come from the momentum-exchange
    probe, reduced on the device into a float64 tensor (or from the immersed body's force);
    dp from the two pressure points, one sparse evaluation per call."""

    def __init__(self, geometry: Geometry):
        self.geometry = geometry
        links = geometry.cylinder_links
        self.probe = None if links is None else ForceProbe.build(geometry.params, links)
        self.history: list = []
        self._points = _pressure_points(geometry)
        self._points_by_layout: dict = {}
        self._force = None

    def _drag_lift(self, solver) -> np.ndarray:
        if self.probe is None:
            return np.asarray(solver.immersed_force)
        if self._force is None:
            torch = solver._torch
            self._force = torch.zeros(self.geometry.params.d, dtype=torch.float64, device=solver.device)
        solver.force_on_device(self.probe, self._force)
        return self._force.cpu().numpy()

    def _pressure_drop(self, solver) -> float:
        p = self.geometry.params
        rows = self._points_by_layout.get(solver.state_layout)
        if rows is None:
            rows = compile_affine(p.lattice, self._points, solver.state_layout, p.solid, p.wall_velocity)
            self._points_by_layout[solver.state_layout] = rows
        front, back = 1.0 + solver.evaluate_affine(rows)
        return CS2 * (front - back) * self.geometry.velocity_scale**2

    def __call__(self, solver) -> np.ndarray:
        g = self.geometry
        coefficients = self._drag_lift(solver) / (0.5 * g.u_mean**2 * g.n)
        values = np.append(coefficients, self._pressure_drop(solver))
        self.history.append((solver.time_step, *values))
        return values


class SteadyResult(NamedTuple):
    n: int
    cd: float
    cl: float
    dp: float
    steps: int


class PeriodicResult(NamedTuple):
    n: int
    cd_max: float
    cl_max: float
    strouhal: float
    steps: int
    history: np.ndarray


def _observed_run(test: str, n: int, backend: str, dtype: str, options: dict):
    """Geometry, a solver started from the inflow profile, and an observer."""
    geometry = build(test, n, dtype=dtype, **options)
    solver = make_solver(geometry.params, backend=backend)
    geometry.start(solver)
    return geometry, solver, Observer(geometry)


def run_steady(n: int = 20, backend: str = "hip-aa", dtype: str = "float64", **options) -> SteadyResult:
    """Test 2D-1 to a steady state, checked once per convective time D / U_mean (to 1e-7 in
    float64, 1e-5 in float32)."""
    geometry, solver, observe = _observed_run("2D-1", n, backend, dtype, options)
    convective = even(n / geometry.u_mean)
    steps = run_until_steady(solver, observe, interval=convective, tol={"float64": 1e-7, "float32": 1e-5}[dtype],
                             max_steps=5_000_000)
    return SteadyResult(n, *map(float, observe(solver)), steps)


def lift_period(t: np.ndarray, cl: np.ndarray) -> float:
    """Mean period of C_L(t) in steps between its upward zero crossings, each placed by
    linear interpolation between the two samples around it."""
    up = np.flatnonzero((cl[:-1] < 0) & (cl[1:] >= 0))
    if up.size < 2:
        raise ValueError("need at least two upward zero crossings of C_L to measure a period")
    fraction = cl[up] / (cl[up] - cl[up + 1])
    crossings = t[up] + fraction * (t[up + 1] - t[up])
    return float(np.diff(crossings).mean())


def run_periodic(n: int = 20, backend: str = "hip-aa", dtype: str = "float32", convective_times: float = 150.0,
                 samples_per_time: int = 40, **options) -> PeriodicResult:
    """Test 2D-2: ``convective_times`` D / U_mean with ``samples_per_time`` samples per
    convective time; C_D, C_L maxima and the Strouhal number from the last quarter."""
    geometry, solver, observe = _observed_run("2D-2", n, backend, dtype, options)
    convective = n / geometry.u_mean
    every = even(convective / samples_per_time)
    total = every * round(convective_times * convective / every)
    solver.run(total, every=every, callback=observe)
    history = np.asarray(observe.history)
    t, cd, cl = history[len(history) * 3 // 4 :, :3].T
    strouhal = n / (lift_period(t, cl) * geometry.u_mean)
    return PeriodicResult(n, float(cd.max()), float(cl.max()), strouhal, total, history)


def main(argv=None) -> None:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--test", default="2D-1", choices=tuple(TESTS))
    parser.add_argument("--sizes", type=int, nargs="+", default=[20, 40])
    parser.add_argument("--method", default="bouzidi", choices=METHODS)
    parser.add_argument("--u-max", type=float, default=None, help="peak inflow, lattice units")
    parser.add_argument("--channel", type=float, nargs="*", default=None, metavar="RE",
                        help="run the 880 x 164 drag channel at these Reynolds numbers instead (default 20 100)")
    parser.add_argument("--steps", type=int, default=50_000, help="steps of a --channel run")
    add_backend_arguments(parser)
    args = parser.parse_args(argv)
    if args.channel is not None:
        print(f"880 x 164 Bouzidi drag channel, {args.collision}, backend = {args.backend}, {args.dtype}, "
              f"drag and lift every step")
        for re in args.channel or [20.0, 100.0]:
            r = run_drag_channel(re, args.steps, args.backend, args.dtype)
            print(f"Re = {re:5.0f}: {r.steps} steps in {r.seconds:.2f} s = {r.mlups:.0f} MLUPS; "
                  f"C_D = {r.cd:.4f}  C_L = {r.cl:+.5f} (mean of the second half)", flush=True)
        return
    print(f"Schaefer-Turek {args.test}, {args.method}, {args.collision}, backend = {args.backend}")
    print("accepted: " + ", ".join(f"{k} in [{lo}, {hi}]" for k, (lo, hi) in INTERVALS[args.test].items()))
    for n in args.sizes:
        options = {"collision": args.collision, "method": args.method, "u_max": args.u_max}
        if args.test == "2D-1":
            r = run_steady(n, args.backend, args.dtype, **options)
            print(f"N = {n:3d}: C_D = {r.cd:.4f}  C_L = {r.cl:.5f}  dp = {r.dp:.4f}  ({r.steps} steps)")
        else:
            r = run_periodic(n, args.backend, args.dtype, **options)
            print(f"N = {n:3d}: C_D,max = {r.cd_max:.4f}  C_L,max = {r.cl_max:.4f}  St = {r.strouhal:.4f}  "
                  f"({r.steps} steps)")


if __name__ == "__main__":
    main()
