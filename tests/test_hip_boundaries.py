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


def _ibm_open_channel(dtype):
    """IBM cylinder in a 64 x 32 channel with a parabolic Zou-He inlet and a pressure outlet."""
    from lbm.boundaries import channel_walls
    from lbm.geometry import circle
    from lbm.ibm import ImmersedBoundary
    from lbm.lattice import D2Q9
    from lbm.open_boundaries import open_channel

    nx, ny = 64, 32
    y = np.arange(ny) - 0.5
    inflow = np.zeros((2, ny))
    inflow[0] = np.where((y > 0) & (y < ny - 2), 4 * 0.04 * y * (ny - 2 - y) / (ny - 2) ** 2, 0.0)
    solid, rules = open_channel(D2Q9, channel_walls(nx, ny), inflow)
    body = ImmersedBoundary.from_ball(circle(19.47, 16.31, 4.0), spacing=1.0, iterations=3)
    return SimulationParams(nx=nx, ny=ny, tau=0.6, collision="trt", solid=solid, dtype=dtype, immersed=body,
                            link_rules=rules)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("case", ["walls-lid", "zou-he"])
def test_forcing_inside_the_pull_step_equals_forcing_then_step(case, dtype):
    """hip-pull runs the IBM forcing inside its step (lbm_ibm_pull_step) when the workgroup
    holds a lane per stencil node, each stencil node's lane summing its moment rows from the
    populations it pulls: against the rows gathered from memory and against the forcing
    launch followed by the step (with the link rules inside it), populations, field and
    body force after 31 steps (graph replays and one eager step) are the same bits."""
    from dataclasses import replace

    params = ibm_case(2, dtype) if case == "walls-lid" else _ibm_open_channel(dtype)
    runs = []
    for mode in ("rows in lanes", "rows gathered", "two launches"):
        s = make_solver(params, backend="hip-pull")
        forcing = s._forcing_in_step()
        assert forcing is not None and forcing.rows_in_lanes
        if mode == "rows gathered":
            s._forcing_in_step = lambda record=False, f=replace(forcing, rows_in_lanes=False): f
        if mode == "two launches":
            s._forcing_in_step = lambda record=False: None
        inside = mode != "two launches"
        calls = s._step_calls(s._stream)[0]
        assert (calls[0][0] == "lbm_ibm_pull_step") == inside and len(calls) == (1 if inside else 2)
        s.initialize(*random_state(params, 2))
        s.run(31)
        runs.append((s.populations(), s._field.cpu().numpy(), s.immersed_force))
    for other in runs[1:]:
        for a, b in zip(runs[0], other, strict=True):
            assert np.array_equal(a, b)
    assert np.any(runs[0][1] != 0)


def test_body_at_rest_in_fluid_at_rest_feels_no_force():
    from lbm import SimulationParams
    from lbm.geometry import circle
    from lbm.ibm import ImmersedBoundary

    params = SimulationParams(nx=24, ny=24, tau=0.8, dtype="float64",
                              immersed=ImmersedBoundary.from_ball(circle(11.5, 11.5, 4.0)))
    solver = make_solver(params, backend="hip-aa")
    solver.run(10)
    np.testing.assert_allclose(solver.immersed_force, 0.0, atol=1e-16)


def rule_sets():
    """Rule sets of the case geometries: a D2Q9 channel with a Bouzidi cylinder and Zou-He
    ends (under 4096 rules: the one-workgroup two-phase pass; over: evaluate + write) and a
    D3Q19 duct with a Bouzidi sphere."""
    from lbm.cases.cylinder_benchmark import drag_channel
    from lbm.cases.sphere_drag import duct_sphere

    small = drag_channel(nx=220, ny=82, radius=10.0, centre=(50.0, 40.0))[0]
    large = drag_channel(nx=1024, ny=600, radius=40.0, centre=(256.0, 300.0))[0]
    duct = duct_sphere(nx=64, ny=30, nz=30, radius=5.0, centre=(20.0, 15.0, 15.0))[0]
    return {"d2q9-small": small, "d2q9-large": large, "d3q19-duct": duct}


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("layout", ["pull", "aa_odd", "aa_even"])
def test_one_launch_pre_pass_equals_the_two_phase_pass(layout, dtype):
    """lbm_rules_apply_direct (evaluate and write each rule in one launch) leaves the state
    bit for bit as lbm_rules_apply (every value before any write) on sets in which no rule
    reads another rule's target, also with per-row widths against lbm_rules_apply_widths."""
    import torch

    from lbm.backends._native import DTYPE_CODES, library, pointer
    from lbm.backends.hip_solver import _AffinePack
    from lbm.link_rules import compile_rules, reads_other_targets

    lib = library()
    code = DTYPE_CODES[dtype]
    tdt = getattr(torch, dtype)
    sets = rule_sets()
    assert len(sets["d2q9-small"].link_rules) <= 4096 < len(sets["d2q9-large"].link_rules)
    for name, params in sets.items():
        compiled = compile_rules(params.lattice, params.link_rules, layout, params.solid, params.wall_velocity)
        assert not reads_other_targets(compiled.target, compiled.values.slots, compiled.values.coeff), name
        pack = _AffinePack(torch, compiled.values, np.dtype(dtype), "cuda")
        target = torch.as_tensor(np.asarray(compiled.target, dtype=np.int64), device="cuda")
        scratch = torch.empty(pack.n, dtype=tdt, device="cuda")
        gen = torch.Generator(device="cuda").manual_seed(7)
        f = torch.rand((params.lattice.q, params.n_nodes), generator=gen, dtype=tdt, device="cuda")
        two, one = f.clone(), f.clone()
        lib.call("lbm_rules_apply", code, pointer(two), pointer(pack.slot), pointer(pack.coeff),
                 pointer(pack.constant), pointer(target), pack.n, pack.m, pointer(scratch), None)
        lib.call("lbm_rules_apply_direct", code, pointer(one), pointer(pack.slot), pointer(pack.coeff),
                 pointer(pack.constant), None, pointer(target), pack.n, pack.m, None)
        torch.cuda.synchronize()
        assert not torch.equal(two, f), name  # the rules changed the state
        assert torch.equal(one, two), name
        # the same set with every row declared one term narrower than the arrays' width
        width = torch.full((pack.n,), pack.m, dtype=torch.int32, device="cuda")
        wide = _AffinePack(torch, type(compiled.values)(
            np.pad(compiled.values.slots, ((0, 0), (0, 1))), np.pad(compiled.values.coeff, ((0, 0), (0, 1))),
            compiled.values.constant), np.dtype(dtype), "cuda")
        two, one = f.clone(), f.clone()
        lib.call("lbm_rules_apply_widths", code, pointer(two), pointer(wide.slot), pointer(wide.coeff),
                 pointer(wide.constant), pointer(target), pointer(width), wide.n, wide.m, pointer(scratch), None)
        lib.call("lbm_rules_apply_direct", code, pointer(one), pointer(wide.slot), pointer(wide.coeff),
                 pointer(wide.constant), pointer(width), pointer(target), wide.n, wide.m, None)
        torch.cuda.synchronize()
        assert torch.equal(one, two), name


@pytest.mark.parametrize(("nf", "fm", "nr", "rm", "d"), [
    (300, 4, 1364, 6, 2), (1024, 8, 2048, 8, 3), (1, 1, 1, 1, 2), (0, 3, 700, 5, 2), (500, 3, 0, 4, 2),
    (300, 9, 1364, 6, 2), (1025, 4, 100, 4, 2), (100, 4, 2049, 4, 3), (2048, 12, 4096, 16, 3),
])
def test_fused_force_and_pre_pass_equal_the_separate_launches(nf, fm, nr, rm, d):
    """lbm_affine_force_rules (the force of one step and the rule pre-pass of the next in one
    workgroup) gives the same bits as lbm_affine_force followed by lbm_rules_apply: padded
    rows, rows of up to 16 terms, one and two rules per lane, sizes at the workgroup's limits
    and empty row sets."""
    import torch

    from lbm.backends._native import library

    rng = np.random.default_rng(nf * 7 + nr)
    n = 9 * 4096
    f0 = torch.as_tensor(rng.uniform(0.5, 1.5, n).astype(np.float32), device="cuda")

    def rows(count, m):
        slot = rng.integers(0, n, (m, max(count, 1))).astype(np.int64)
        coeff = rng.normal(size=(m, max(count, 1))).astype(np.float32)
        width = rng.integers(1, m + 1, max(count, 1))
        pad = np.arange(m)[:, None] >= width[None, :]
        coeff[pad], slot[pad] = 0.0, 0
        constant = rng.normal(size=max(count, 1)).astype(np.float32)
        return [torch.as_tensor(a, device="cuda") for a in (slot, coeff, constant)]

    fs, fc, fk = rows(nf, fm)
    weight = torch.as_tensor(rng.normal(size=(max(nf, 1), d)), device="cuda")
    rs, rc, rk = rows(nr, rm)
    target = torch.as_tensor(rng.permutation(n)[: max(nr, 1)].astype(np.int64), device="cuda")
    lib = library()
    p = lambda t: t.data_ptr()  # noqa: E731
    fused, separate = f0.clone(), f0.clone()
    out_fused = torch.full((3,), 7.0, dtype=torch.float64, device="cuda")
    out_separate = torch.full((3,), 7.0, dtype=torch.float64, device="cuda")
    lib.call("lbm_affine_force_rules", 0, p(fused), p(fs), p(fc), p(fk), p(weight), nf, fm, d, p(out_fused), p(rs),
             p(rc), p(rk), p(target), nr, rm, None)
    lib.call("lbm_affine_force", 0, p(separate), p(fs), p(fc), p(fk), p(weight), nf, fm, d, p(out_separate), None, None)
    lib.call("lbm_rules_apply", 0, p(separate), p(rs), p(rc), p(rk), p(target), nr, rm, None, None)
    torch.cuda.synchronize()
    assert torch.equal(fused, separate)
    assert torch.equal(out_fused[:d], out_separate[:d])
    if nr:
        assert not torch.equal(fused, f0)


def old_pull_steps(solver, steps):
    """The pre-pass form of the two-lattice step: lbm_rules_apply(_direct) on the source,
    then lbm_pull_step (the launches before the rules moved into the step)."""
    from lbm.backends._native import pointer

    L, D = solver._lattice_code, solver._dtype_code
    for _ in range(steps):
        src, dst = pointer(solver._f), pointer(solver._f_next)
        calls = solver._pre_pass_calls("pull", src, solver._stream)
        calls.append(("lbm_pull_step", (L, D, src, dst, pointer(solver._kind), solver._wall_ptr,
                                        pointer(solver._field), pointer(solver._body), *solver._step_args())))
        solver._lib.call_all(calls)
        solver._f, solver._f_next = solver._f_next, solver._f


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("d", [2, 3])
def test_rules_evaluated_in_the_step_equal_the_pre_pass_form(d, dtype):
    """lbm_pull_step_links (each rule row evaluated by the lane that consumes it) against
    the rule pre-pass followed by lbm_pull_step: the same bits after 25 steps, for Bouzidi
    bodies with moving walls and for Zou-He inlet/outlet channels."""
    from dataclasses import replace

    for params in (bouzidi_case(d, dtype)[0], replace(open_case(d)[0], dtype=dtype)):
        a, b = make_solver(params, backend="hip-pull"), make_solver(params, backend="hip-pull", rules_in_step=False)
        assert a._links is not None and a._links.n == len(params.link_rules) and b._links is None
        for s in (a, b):
            s.initialize(*random_state(params, 4))
        a.run(25)
        old_pull_steps(b, 25)
        np.testing.assert_array_equal(a.populations(), b.populations())


@pytest.mark.parametrize("steps", [6, 7])
def test_step_layer_force_history_of_a_large_probe_equals_the_separate_launches(steps):
    """More than 2048 probe rows: the step's force layer writes partial sums per group,
    lbm_force_totals adds them; every row equals lbm_affine_force on the same state."""
    import torch

    from lbm.geometry import sphere, voxelize, wall_links
    from lbm.lattice import D3Q19
    from lbm.link_rules import bouzidi_rules

    shape = (28, 28, 36)
    body = sphere(17.3, 13.6, 14.2, 7.4)
    solid = voxelize(body, shape)
    links = wall_links(D3Q19, body, solid)
    params = SimulationParams(nx=36, ny=28, nz=28, tau=0.7, collision="trt", solid=solid, dtype="float32",
                              link_rules=bouzidi_rules(D3Q19, links, solid), force=(1e-5, 0.0, 0.0))
    probe = ForceProbe.build(params, links)
    assert len(links) > 2048
    fast, slow = make_solver(params, backend="hip-pull"), make_solver(params, backend="hip-pull")
    for s in (fast, slow):
        s.initialize(*random_state(params, 1))
    recorded = fast.record_forces(probe, steps)
    expected = torch.zeros((steps, 3), dtype=torch.float64, device="cuda")
    for k in range(steps):
        slow.run(1, check_every=1)
        slow.force_on_device(probe, expected[k])
    assert np.array_equal(recorded, expected.cpu().numpy())
    assert np.array_equal(fast.populations(), slow.populations())


def large_sphere_case(dtype):
    """A sphere of radius 5.3 (about 400 markers, about 2,500 stencil nodes): its forcing
    intermediates do not fit one workgroup's LDS, so lbm_ibm_force takes its
    many-workgroup launches."""
    from lbm import SimulationParams
    from lbm.boundaries import channel_walls
    from lbm.geometry import sphere
    from lbm.ibm import ImmersedBoundary

    solid = channel_walls(30, 24, 22)
    body = ImmersedBoundary.from_ball(sphere(14.3, 11.6, 10.8, 5.3))
    return SimulationParams(nx=30, ny=24, nz=22, tau=0.7, collision="trt", force=(1e-5, -2e-6, 3e-6), solid=solid,
                            dtype=dtype, immersed=body)


def test_many_workgroup_ibm_of_a_large_sphere_is_bit_identical_to_the_oracle():
    """Pull layout, float64: 10 steps of the oracle and of hip-pull, whose forcing runs as
    the many-workgroup launches (the intermediates need more LDS than a workgroup has)."""
    import ctypes

    from lbm.backends._native import library

    params = large_sphere_case("float64")
    ibm_nodes = make_solver(params, backend="hip-pull")._ibm
    size = ctypes.c_int64()
    library().raw("lbm_ibm_scratch_size")(1, 1, ibm_nodes.n_nodes, ibm_nodes.n_markers, ctypes.byref(size))
    assert size.value > 160 * 1024  # larger than one workgroup's LDS
    ref, gpu = both(params, "hip-pull", steps=10)
    np.testing.assert_array_equal(gpu.populations()[:, params.fluid], ref.populations()[:, params.fluid])
    np.testing.assert_array_equal(gpu.immersed_force, ref.immersed_force)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("d", [2, 3])
def test_many_workgroup_ibm_equals_the_one_workgroup_kernel(d, dtype):
    """lbm_ibm_force_multi against lbm_ibm_force (one workgroup, tables staged in LDS) on the
    same state: the same field and body force, bit for bit."""
    import torch

    from lbm.backends._native import pointer

    params = ibm_case(d, dtype)
    s = make_solver(params, backend="hip-pull")
    s.initialize(*random_state(params, 5))
    s.run(3)
    calls = s._ibm_calls("pull", pointer(s._f), s._stream)
    assert len(calls) == 1 and calls[0][0] == "lbm_ibm_force"
    name, args = calls[0]
    outs = []
    for entry in ("lbm_ibm_force", "lbm_ibm_force_multi"):
        field = torch.zeros_like(s._field)
        body = torch.zeros_like(s._body_force)
        scratch = torch.full_like(s._ibm_scratch, float("nan"))
        a = list(args)
        a[20], a[21], a[22] = pointer(field), pointer(body), pointer(scratch)
        if entry == "lbm_ibm_force_multi":
            del a[23]  # no record argument
        s._call(entry, *a)
        torch.cuda.synchronize()
        outs.append((field.cpu().numpy(), body.cpu().numpy()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert np.abs(outs[0][0]).max() > 0


def test_large_lattices_keep_the_rule_pre_pass_and_both_forms_agree():
    """Above RULES_IN_STEP_MAX_NODES the pull solver applies its rules in a pre-pass by
    default; either form, asked for explicitly, gives the same bits."""
    from lbm.backends.hip_solver import RULES_IN_STEP_MAX_NODES
    from lbm.cases.sphere_drag import duct_sphere

    params, _, inflow = duct_sphere(nx=256, ny=128, nz=136, radius=8.0, centre=(64.0, 64.0, 68.0))
    assert params.n_nodes > RULES_IN_STEP_MAX_NODES
    default = make_solver(params, backend="hip-pull")
    assert not default.rules_in_step and default._links is None
    in_step = make_solver(params, backend="hip-pull", rules_in_step=True)
    ux = np.where(params.solid, 0.0, inflow[0][:, :, None])
    for s in (default, in_step):
        s.initialize(1.0, ux, 0.0, 0.0)
        s.run(6)
    np.testing.assert_array_equal(default.populations(), in_step.populations())
