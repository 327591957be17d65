"""Link rules (Bouzidi, Zou-He), the momentum-exchange probe and the fused IBM kernel on the GPU,
against the NumPy oracle (tests/test_geometry_link_rules.py:121-246, test_open_boundaries.py:63-72
and test_ibm.py:59-101 of the reference). Each case exercises the rule compiler for the layout the
backend reads (pull; AA odd and even)."""

import numpy as np
import pytest

from lbm import SimulationParams, make_solver
from lbm.forces import ForceProbe
from tests.helpers import fluid_diff, random_state
from tests.test_golden import bouzidi_case, ibm_case, open_case

pytestmark = pytest.mark.gpu
BACKENDS = ["hip-pull", "hip-aa", "hip-lds"]


def both(params, backend, seed=0, steps=40):
    from oracle import ReferenceSolver

    ref = ReferenceSolver(params)
    gpu = make_solver(params, backend=backend)
    for s in (ref, gpu):
        s.initialize(*random_state(params, seed))
        s.run(steps)
    return ref, gpu


def backends_for(d):
    return BACKENDS if d == 2 else BACKENDS[:2]


CASES = [(b, d) for d in (2, 3) for b in backends_for(d)]


@pytest.mark.parametrize(("backend", "d"), CASES)
@pytest.mark.parametrize(("dtype", "tol"), [("float64", 1e-15), ("float32", 1e-7)])
def test_bouzidi_matches_the_oracle(backend, d, dtype, tol):
    params, _, _ = bouzidi_case(d, dtype)
    ref, gpu = both(params, backend)
    assert fluid_diff(params, ref, gpu) < tol


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("d", [2, 3])
def test_pull_with_link_rules_is_bit_identical_to_the_oracle(d, dtype):
    """Zou-He inlet/outlet and Bouzidi rows evaluated on the device in NumPy's summation
    order: the two-lattice states equal the oracle's bit for bit after 40 steps."""
    for params in (bouzidi_case(d, dtype)[0], open_case(d)[0]):
        if params.dtype != dtype:  # the open-channel case is built in float64
            from dataclasses import replace

            params = replace(params, dtype=dtype)
        ref, gpu = both(params, "hip-pull", seed=2)
        fluid = params.fluid
        np.testing.assert_array_equal(gpu.populations()[:, fluid], ref.populations()[:, fluid])


@pytest.mark.parametrize(("backend", "d"), CASES)
def test_force_probe_reads_every_layout_the_same(backend, d):
    params, links, _ = bouzidi_case(d, "float64")
    probe = ForceProbe.build(params, links)
    ref, gpu = both(params, backend)
    np.testing.assert_allclose(probe.evaluate(gpu), probe.evaluate(ref), rtol=1e-12, atol=1e-16)


@pytest.mark.parametrize(("backend", "d"), CASES)
def test_force_on_device_matches_the_host_evaluation(backend, d):
    import torch

    params, links, _ = bouzidi_case(d, "float64")
    probe = ForceProbe.build(params, links)
    _, gpu = both(params, backend)
    out = torch.zeros(d, dtype=torch.float64, device="cuda")
    gpu.force_on_device(probe, out)
    np.testing.assert_allclose(out.cpu().numpy(), probe.evaluate(gpu), rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("backend", ["hip-pull", "hip-aa"])
def test_force_of_a_large_probe_is_reduced_by_many_groups_deterministically(backend):
    """A sphere probe with more rows than one workgroup reduces (partial sums per group,
    then a fixed-order total): equal to the host evaluation, and bit for bit the same on
    every call."""
    import torch

    from lbm.geometry import sphere, voxelize, wall_links
    from lbm.lattice import D3Q19
    from lbm.link_rules import bouzidi_rules

    shape = (28, 28, 36)
    body = sphere(17.3, 13.6, 14.2, 7.4)
    solid = voxelize(body, shape)
    links = wall_links(D3Q19, body, solid)
    params = SimulationParams(nx=36, ny=28, nz=28, tau=0.7, collision="trt", solid=solid, dtype="float64",
                              link_rules=bouzidi_rules(D3Q19, links, solid), force=(1e-5, 0.0, 0.0))
    probe = ForceProbe.build(params, links)
    assert len(links) > 2048  # 4 groups of 1024 rows
    _, gpu = both(params, backend)
    outs = [torch.zeros(3, dtype=torch.float64, device="cuda") for _ in range(3)]
    for out in outs:
        gpu.force_on_device(probe, out)
    host = probe.evaluate(gpu)
    np.testing.assert_allclose(outs[0].cpu().numpy(), host, rtol=1e-11, atol=1e-15 * np.abs(host).max())
    for out in outs[1:]:
        assert np.array_equal(out.cpu().numpy(), outs[0].cpu().numpy())


@pytest.mark.parametrize(("backend", "d"), CASES)
def test_open_boundaries_match_the_oracle(backend, d):
    params, _ = open_case(d)
    ref, gpu = both(params, backend, seed=1)
    assert fluid_diff(params, ref, gpu) < 1e-15


@pytest.mark.parametrize(("backend", "d"), CASES)
@pytest.mark.parametrize(("dtype", "tol"), [("float64", 1e-15), ("float32", 1e-7)])
def test_ibm_matches_the_oracle(backend, d, dtype, tol):
    params = ibm_case(d, dtype)
    ref, gpu = both(params, backend)
    assert fluid_diff(params, ref, gpu) < tol
    np.testing.assert_allclose(gpu.immersed_force, ref.immersed_force, rtol=10 * tol, atol=1e-18)


@pytest.mark.parametrize("d", [2, 3])
def test_ibm_pull_is_bit_identical_to_the_oracle(d):
    """Pull layout: same moment rows, same pairwise and bincount orders -> same bits."""
    params = ibm_case(d, "float64")
    ref, gpu = both(params, "hip-pull")
    np.testing.assert_array_equal(gpu.populations()[:, params.fluid], ref.populations()[:, params.fluid])
    np.testing.assert_array_equal(gpu.immersed_force, ref.immersed_force)


def test_body_at_rest_in_fluid_at_rest_feels_no_force():
    from lbm import SimulationParams
    from lbm.geometry import circle
    from lbm.ibm import ImmersedBoundary

    params = SimulationParams(nx=24, ny=24, tau=0.8, dtype="float64",
                              immersed=ImmersedBoundary.from_ball(circle(11.5, 11.5, 4.0)))
    solver = make_solver(params, backend="hip-aa")
    solver.run(10)
    np.testing.assert_allclose(solver.immersed_force, 0.0, atol=1e-16)
