"""Exact-solution validation (tests/validation/test_poiseuille.py, test_taylor_green.py,
test_beltrami.py of the reference): TRT Poiseuille to round-off, BGK slip closed form,
second-order Taylor-Green / ABC convergence, on the oracle (CPU) and the HIP backends (GPU)."""

import math

import numpy as np
import pytest

from lbm.cases.beltrami import run_beltrami
from lbm.cases.poiseuille import BGK_EXACT_TAU, run_poiseuille
from lbm.cases.taylor_green import convergence_order, run_taylor_green
from oracle import ReferenceSolver

ROUND_OFF = 1e-10
GPU = [pytest.param(b, marks=pytest.mark.gpu) for b in ("hip-pull", "hip-aa")]
ORACLE = pytest.param("numpy")


def factory(backend):
    return ReferenceSolver if backend == "numpy" else None


@pytest.mark.parametrize("backend", [ORACLE, *GPU])
@pytest.mark.parametrize("tau", [0.6, 2.0])
def test_trt_poiseuille_is_exact_for_any_tau(backend, tau):
    result = run_poiseuille(tau=tau, collision="trt", backend=backend, solver_factory=factory(backend))
    assert result.l2_error < ROUND_OFF


@pytest.mark.parametrize("backend", [ORACLE, *GPU])
def test_bgk_is_exact_only_at_its_magic_tau(backend):
    result = run_poiseuille(tau=BGK_EXACT_TAU, collision="bgk", backend=backend, solver_factory=factory(backend))
    assert result.l2_error < ROUND_OFF


@pytest.mark.parametrize("tau", [0.6, 1.5])
def test_bgk_error_is_the_closed_form_slip(tau):
    height, umax = 16, 0.01
    result = run_poiseuille(height=height, tau=tau, collision="bgk", max_velocity=umax, solver_factory=ReferenceSolver)
    nu = (tau - 0.5) / 3.0
    force = 8.0 * nu * umax / height**2
    slip = (2.0 / 3.0) * ((tau - 0.5) ** 2 - 3.0 / 16.0) * force / nu
    np.testing.assert_allclose(result.u - result.u_exact, slip, rtol=1e-9, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
@pytest.mark.parametrize("tau", [0.6, 2.0])
def test_3d_slab_trt_is_exact_on_the_gpu(backend, tau):
    assert run_poiseuille(tau=tau, collision="trt", backend=backend, nz=4).l2_error < ROUND_OFF


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
def test_float32_trt_poiseuille_stays_at_single_precision(backend):
    assert run_poiseuille(tau=0.8, collision="trt", backend=backend, dtype="float32").l2_error < 1e-4


def test_taylor_green_converges_at_second_order_on_the_oracle():
    results = [run_taylor_green(n=n, solver_factory=ReferenceSolver) for n in (16, 24, 32)]
    assert all(r.decay_rate_error < 2e-2 for r in results)
    assert all(1.9 < o < 2.1 for o in convergence_order(results))


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa", "hip-lds"])
def test_taylor_green_converges_at_second_order_on_the_gpu(backend):
    results = [run_taylor_green(n=n, backend=backend) for n in (32, 64, 128)]
    assert results[0].decay_rate_error < 1e-2
    assert all(1.9 < o < 2.1 for o in convergence_order(results)), convergence_order(results)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
@pytest.mark.parametrize("collision", ["bgk", "trt"])
def test_abc_decay_rate_matches_the_lattice_viscosity(backend, collision):
    assert run_beltrami(n=32, collision=collision, backend=backend).decay_rate_error < 1e-2


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
def test_abc_converges_at_second_order_on_the_gpu(backend):
    results = [run_beltrami(n=n, backend=backend) for n in (16, 32, 64, 128)]
    orders = convergence_order(results)
    assert all(1.9 < o < 2.1 for o in orders), orders


@pytest.mark.gpu
def test_abc_float32_matches_float64_accuracy():
    single = run_beltrami(n=32, backend="hip-aa", dtype="float32")
    double = run_beltrami(n=32, backend="hip-aa", dtype="float64")
    assert single.l2_error == pytest.approx(double.l2_error, rel=1e-2)


def test_abc_converges_on_the_oracle():
    results = [run_beltrami(n=n, solver_factory=ReferenceSolver) for n in (16, 24)]
    assert results[1].l2_error < results[0].l2_error
    assert 1.8 < convergence_order(results)[0] < 2.2


def periodic_disk_params(rules: bool):
    """A disk in a fully periodic box pushed by a uniform body force (the reference's
    tests/test_geometry_link_rules.py:195-209)."""
    from lbm import SimulationParams
    from lbm.geometry import circle, voxelize, wall_links
    from lbm.lattice import D2Q9
    from lbm.link_rules import bouzidi_rules

    size, body = 24, circle(11.5, 11.5, 3.3)
    solid = voxelize(body, (size, size))
    links = wall_links(D2Q9, body, solid)
    params = SimulationParams(nx=size, ny=size, tau=1.0, collision="trt", force=(1e-6, 0.0), solid=solid,
                              dtype="float64", link_rules=bouzidi_rules(D2Q9, links, solid) if rules else None)
    return params, links


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
@pytest.mark.parametrize("rules", [False, True], ids=["plain", "bouzidi"])
def test_steady_momentum_exchange_balances_the_body_force(backend, rules):
    """At steady state the momentum exchange on the disk equals the body force on all the
    fluid, to 1e-9 relative (tests/test_geometry_link_rules.py:213-221), read on the device
    (force_on_device) and on the host (ForceProbe.evaluate)."""
    import torch

    from lbm import make_solver
    from lbm.forces import ForceProbe

    params, links = periodic_disk_params(rules)
    solver = make_solver(params, backend=backend)
    solver.run(6000)
    probe = ForceProbe.build(params, links)
    total = params.force[0] * params.fluid.sum()
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    solver.force_on_device(probe, out)
    for force in (probe.evaluate(solver), out.cpu().numpy()):
        assert force[0] == pytest.approx(total, rel=1e-9)
        assert abs(force[1]) < 1e-12 * total


@pytest.mark.gpu
def test_cavity_re1000_matches_ghia():
    """Ghia, Ghia & Shin Re = 1000 at 128^2 (tests/validation/test_cavity.py:26-29): largest
    centreline deviation below 0.025 lid velocities (the reference measured 0.0148)."""
    from lbm.cases.lid_driven_cavity import run_cavity

    assert run_cavity(reynolds=1000, n=128, backend="hip-aa", dtype="float32").max_error < 0.025


@pytest.mark.gpu
def test_ibm_sphere_is_too_draggy_at_coarse_resolution():
    """The immersed-boundary sphere at N = 10, Re = 20 (tests/validation/test_sphere_drag.py:
    23-25): C_D 10-25 % above Clift-Grace-Weber, as the reference measured (3.112 against
    2.735), the diffuse interface widening the body."""
    from lbm.cases.sphere_drag import run_sphere_drag

    ibm = run_sphere_drag(10, 20.0, "ibm")
    assert 0.10 < ibm.cd / ibm.reference_cgw - 1 < 0.25
