"""The host-side API a caller of the reference uses, unchanged: SimulationParams validation
and normalisation (tests/test_params.py of the reference), the Solver run loop
(tests/test_solver.py) and the diagnostics (tests/test_diagnostics.py). The run-loop tests
use a counting backend, so they need no GPU; the HipSolver's own behaviour is checked on
the GPU at the end."""

import math
import warnings

import numpy as np
import pytest

from lbm import SimulationParams
from lbm.diagnostics import MacroscopicFields, kinetic_energy, vorticity
from lbm.lattice import D2Q9, D3Q19
from lbm.params import StabilityWarning
from lbm.solver import SimulationDiverged, Solver

# ---------------------------------------------------------------- parameters


def test_defaults_are_fully_periodic_fluid():
    p = SimulationParams(nx=8, ny=6, tau=0.8)
    assert p.shape == (6, 8) and p.d == 2 and p.lattice is D2Q9
    assert not p.solid.any() and p.fluid.all()
    assert p.force == (0.0, 0.0) and p.collision == "bgk" and p.dtype == "float32"


def test_viscosity_and_relaxation_rates():
    p = SimulationParams(nx=4, ny=4, tau=0.8)
    assert p.viscosity == pytest.approx(0.1)
    assert p.relaxation_rates == (1.25, 1.25)  # BGK: one rate
    for magic in (3 / 16, 1 / 4, 1 / 12):
        q = SimulationParams(nx=4, ny=4, tau=0.8, collision="trt", magic=magic)
        wp, wm = q.relaxation_rates
        assert wp == 1.25
        assert (1 / wp - 0.5) * (1 / wm - 0.5) == pytest.approx(magic)


@pytest.mark.parametrize("tau", [0.5, 0.3, -1.0])
def test_tau_at_or_below_half_is_rejected(tau):
    with pytest.raises(ValueError, match="tau"):
        SimulationParams(nx=4, ny=4, tau=tau)


def test_tau_close_to_half_warns():
    with pytest.warns(StabilityWarning, match="close to 0.5"):
        SimulationParams(nx=4, ny=4, tau=0.505)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        SimulationParams(nx=4, ny=4, tau=0.6)


@pytest.mark.parametrize(("kwargs", "match"), [
    ({"collision": "mrt"}, "collision"), ({"collision": "trt", "magic": 0.0}, "magic"),
    ({"dtype": "float16"}, "dtype"), ({"nx": 0}, "positive"), ({"nz": -2}, "positive"),
    ({"solid": np.zeros((3, 3), dtype=bool)}, "solid mask"),
    ({"wall_velocity": np.zeros((2, 3, 3))}, "wall_velocity"),
    ({"force": (1.0,)}, "components"), ({"lattice": "D3Q27"}, "unknown lattice"),
    ({"lattice": "D3Q19"}, "2D"),
])
def test_invalid_parameters_are_rejected(kwargs, match):
    base = {"nx": 4, "ny": 4, "tau": 0.8} | kwargs
    with pytest.raises(ValueError, match=match):
        SimulationParams(**base)


def test_arrays_are_normalised_read_only_copies():
    solid = np.zeros((4, 5), dtype=bool)
    solid[0] = True
    wall = np.ones((2, 4, 5))
    p = SimulationParams(nx=5, ny=4, tau=0.8, solid=solid, wall_velocity=wall, force=[1, 2])
    solid[1] = True  # the caller's array is copied
    assert p.solid.sum() == 5 and not p.solid.flags.writeable and not p.wall_velocity.flags.writeable
    assert np.all(p.wall_velocity[:, 1:] == 0.0) and np.all(p.wall_velocity[:, 0] == 1.0)  # zeroed on fluid
    assert p.force == (1.0, 2.0) and all(isinstance(f, float) for f in p.force)


def test_3d_uses_d3q19_with_zyx_shape():
    p = SimulationParams(nx=5, ny=4, tau=0.8, nz=3, force=(0, 0, 1e-6))
    assert p.lattice is D3Q19 and p.shape == (3, 4, 5) and p.n_nodes == 60
    assert SimulationParams(nx=5, ny=4, tau=0.8, nz=3, lattice=D3Q19).lattice is D3Q19
    with pytest.raises(ValueError, match="3 components"):
        SimulationParams(nx=5, ny=4, tau=0.8, nz=3, force=(1.0, 0.0))


# ---------------------------------------------------------------- run loop


class Counting(Solver):
    """A backend that only counts: _advance records chunk lengths."""

    def __init__(self, params, multiple=1, diverge_at=None):
        super().__init__(params)
        self.step_multiple = multiple
        self.chunks = []
        self.diverge_at = diverge_at

    def initialize(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        self.fields = self._initial_fields(rho, ux, uy, uz)

    def _advance(self, steps):
        self.chunks.append(steps)

    def _check_health(self):
        if self.diverge_at is not None and self.time_step >= self.diverge_at:
            raise SimulationDiverged(self.time_step, (1, 2), "non-finite populations")

    def populations(self):
        return np.zeros((9, *self.params.shape))

    def evaluate_affine(self, compiled):
        return np.zeros(0)


def test_run_counts_steps_and_fires_callback_every_interval():
    s = Counting(SimulationParams(nx=4, ny=4, tau=0.8))
    seen = []
    s.run(25, every=10, callback=lambda solver: seen.append(solver.time_step), check_every=7)
    assert s.time_step == 25 and sum(s.chunks) == 25
    assert seen == [10, 20, 25]  # every interval, and after the last step
    assert max(s.chunks) <= 7


@pytest.mark.parametrize("kwargs", [{"steps": -1}, {"steps": 4, "every": -2}, {"steps": 4, "check_every": 0}])
def test_run_rejects_negative_counts(kwargs):
    steps = kwargs.pop("steps")
    with pytest.raises(ValueError, match="need steps"):
        Counting(SimulationParams(nx=4, ny=4, tau=0.8)).run(steps, **kwargs)


def test_pairwise_backends_need_even_counts_and_health_stops_the_run():
    s = Counting(SimulationParams(nx=4, ny=4, tau=0.8), multiple=2)
    with pytest.raises(ValueError, match="multiples of 2"):
        s.run(3)
    s.run(6, every=2)
    assert s.chunks == [2, 2, 2]
    bad = Counting(SimulationParams(nx=4, ny=4, tau=0.8), diverge_at=10)
    with pytest.raises(SimulationDiverged, match=r"step 10: non-finite populations at node \(y, x\)=\(1, 2\)"):
        bad.run(30, check_every=5)
    assert bad.time_step == 10


def test_initial_fields_carry_half_the_force_and_refuse_2d_z_velocity():
    p = SimulationParams(nx=3, ny=2, tau=0.8, force=(2e-3, 0.0), dtype="float64")
    s = Counting(p)
    s.initialize(rho=2.0, ux=0.01)
    drho, (ux, uy) = s.fields
    assert np.all(drho == 1.0) and np.allclose(ux, 0.01 + 0.5 * 2e-3 / 2.0) and np.all(uy == 0.0)
    with pytest.raises(ValueError, match="no z velocity"):
        s.initialize(uz=0.1)


# ---------------------------------------------------------------- diagnostics


def test_vorticity_of_2d_solid_body_rotation_is_twice_the_angular_velocity():
    y, x = np.mgrid[0:16, 0:16].astype(float)
    omega = 0.01
    w = vorticity(-omega * (y - 7.5), omega * (x - 7.5))
    assert np.allclose(w, 2 * omega)
    assert np.allclose(vorticity(np.full((5, 5), 0.1), np.zeros((5, 5))), 0.0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_vorticity_of_3d_solid_body_rotation_points_along_the_axis(axis):
    z, y, x = np.mgrid[0:8, 0:8, 0:8].astype(float) - 3.5
    r = [x, y, z]
    omega = np.zeros(3)
    omega[axis] = 0.02
    u = np.cross(omega, np.stack(r), axis=0)  # u = omega x r
    w = vorticity(*u)
    for a in range(3):
        assert np.allclose(w[a], 2 * omega[a])
    with pytest.raises(ValueError, match="2 or 3"):
        vorticity(x)


def test_kinetic_energy_and_macroscopic_fields():
    rho = np.full((2, 3), 2.0)
    ux, uy = np.full((2, 3), 0.1), np.full((2, 3), 0.2)
    assert kinetic_energy(rho, ux, uy) == pytest.approx(0.5 * 2.0 * 0.05 * 6)
    solid = np.zeros((2, 3), dtype=bool)
    solid[0, 0] = True
    f = MacroscopicFields(rho=rho.copy(), u=[ux.copy(), uy.copy()], solid=solid)
    assert f.d == 2 and f.ux is f.u[0] and f.uy is f.u[1]
    assert np.allclose(f.speed, math.sqrt(0.05)) and f.total_mass == pytest.approx(10.0)
    with pytest.raises(AttributeError, match="uz"):
        f.uz
    with pytest.raises(ValueError):
        f.rho[0, 0] = 3.0


# ---------------------------------------------------------------- HipSolver


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
def test_hip_solver_reports_the_requested_velocity_and_returns_host_copies(backend):
    from lbm import make_solver

    p = SimulationParams(nx=12, ny=10, tau=0.8, force=(1e-4, -2e-5), dtype="float64")
    s = make_solver(p, backend=backend)
    s.initialize(1.0, 0.02, -0.01)
    m = s.macroscopic()
    np.testing.assert_allclose(m.ux, 0.02, rtol=1e-12)
    np.testing.assert_allclose(m.uy, -0.01, rtol=1e-12)
    assert not m.rho.flags.writeable
    f = s.populations()
    f[:] = 7.0  # a host copy: the device state is unchanged
    assert not np.any(s.populations() == 7.0)
    with pytest.raises(ValueError, match="no z velocity"):
        s.initialize(uz=0.1)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
def test_reference_benchmark_pattern_drives_the_hip_solver(backend):
    """The reference's timing helpers (benchmarks/bench_mlups.py:99-109) use ``solver.xp``
    and ``solver.run`` only; they run unchanged on the HIP solver, and
    ``_device_populations`` is the natural-order device state (src/lbm/solver.py:89-92)."""
    import time

    import torch

    from lbm import make_solver
    from oracle import ReferenceSolver

    def _synchronize(solver) -> None:
        if solver.xp is not np:
            solver.xp.cuda.get_current_stream().synchronize()

    def _timed_run(solver, steps: int) -> float:
        _synchronize(solver)
        start = time.perf_counter()
        solver.run(steps, check_every=steps)
        _synchronize(solver)
        return time.perf_counter() - start

    p = SimulationParams(nx=24, ny=20, nz=16, tau=0.8, collision="trt", dtype="float64")
    s = make_solver(p, backend=backend)
    rng = np.random.default_rng(0)
    velocity = [0.01 * rng.standard_normal(p.shape) for _ in range(3)]
    s.initialize(1.0, *velocity)
    probe = 2 * s.step_multiple
    assert _timed_run(s, probe) > 0 and s.time_step == probe
    assert s.xp is not np and s.xp.asarray(np.arange(3)).device.type == "cuda"
    assert bool(s.xp.isfinite(s.xp.asarray([1.0, 2.0])).all())

    ref = ReferenceSolver(p)
    ref.initialize(1.0, *velocity)
    ref.run(probe)
    f = s._device_populations()
    assert isinstance(f, torch.Tensor) and f.is_cuda and tuple(f.shape) == (19, *p.shape)
    np.testing.assert_array_equal(f.cpu().numpy(), s.populations())
    np.testing.assert_allclose(f.cpu().numpy(), ref.populations(), rtol=0, atol=1e-15)
