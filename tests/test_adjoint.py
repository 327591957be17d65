"""The hand-written adjoint: gradients of the mean IBM drag w.r.t. a dense body-force field and
the marker positions, checked against float64 torch autograd through a torch re-expression of
the oracle step and against float64 central finite differences of the forward run.

The reference has no adjoint; the checks pin it to the mathematics (autograd of the same
step, finite differences) rather than to reference outputs.
"""

import numpy as np
import pytest
import torch

from lbm import SimulationParams
from lbm.boundaries import channel_walls
from lbm.geometry import circle
from lbm.ibm import ImmersedBoundary, roma, roma_derivative
from lbm.lattice import D2Q9
from lbm.open_boundaries import open_channel
from oracle import ReferenceSolver
from oracle.torch_step import TorchDragModel

STEPS, WINDOW, K = 60, 20, 20


def case(dtype="float64", rules=False, nx=64, ny=32):
    solid = channel_walls(nx, ny)
    link_rules = None
    force = (1e-5, 0.0)
    if rules:
        y = np.arange(ny) - 0.5
        h = ny - 2
        inflow = np.zeros((2, ny))
        inflow[0] = np.where((y > 0) & (y < h), 4 * 0.04 * y * (h - y) / h**2, 0.0)
        solid, link_rules = open_channel(D2Q9, solid, inflow)
        force = (0.0, 0.0)
    body = ImmersedBoundary.from_ball(circle(0.3 * nx + 0.27, ny / 2 + 0.31, 4.0), spacing=1.0, iterations=3)
    return SimulationParams(nx=nx, ny=ny, tau=0.6, collision="trt", force=force, solid=solid, dtype=dtype,
                            immersed=body, link_rules=link_rules)


def init_fields(params):
    return (1.0, np.full(params.shape, 0.03), np.zeros(params.shape))


def oracle_initial_state(params):
    s = ReferenceSolver(params)
    s.initialize(*init_fields(params))
    return s.populations()


# ---------------------------------------------------------------- CPU: the checker itself


def test_roma_derivative_matches_finite_differences():
    r = np.linspace(-1.7, 1.7, 341)
    # phi is C1: its second derivative jumps at |r| = 1/2 and 3/2, where central
    # differences are only first-order accurate, so those points are left out.
    r = r[(np.abs(np.abs(r) - 0.5) > 1e-3) & (np.abs(np.abs(r) - 1.5) > 1e-3)]
    eps = 1e-6
    fd = (roma(r + eps) - roma(r - eps)) / (2 * eps)
    np.testing.assert_allclose(roma_derivative(r), fd, atol=1e-7)


def test_marker_gradient_is_the_chain_rule_of_the_roma_weights():
    from lbm.adjoint import marker_gradient

    body = ImmersedBoundary.from_ball(circle(10.2, 9.7, 4.0), spacing=1.0)
    grad_w = np.random.default_rng(0).standard_normal((len(body.markers), 9))
    base = body.stencil_base()
    g = marker_gradient(body.markers, base, grad_w)
    eps = 1e-6
    for l, a in [(0, 0), (3, 1), (7, 0)]:
        plus, minus = body.markers.copy(), body.markers.copy()
        plus[l, a] += eps
        minus[l, a] -= eps
        wp = body.with_markers(plus).stencil((20, 21))[1]
        wm = body.with_markers(minus).stencil((20, 21))[1]
        fd = np.sum(grad_w * (wp - wm)) / (2 * eps)
        assert g[l, a] == pytest.approx(fd, rel=1e-6, abs=1e-10)


def test_transposed_rows_are_the_transpose():
    from lbm.adjoint import transpose_rows

    rng = np.random.default_rng(1)
    slots = rng.integers(0, 30, size=(12, 4))
    coeff = rng.standard_normal((12, 4)) * (rng.random((12, 4)) > 0.3)
    uslot, start, row, c = transpose_rows(slots, coeff)
    dense = np.zeros((12, 30))
    for r in range(12):
        for t in range(4):
            dense[r, slots[r, t]] += coeff[r, t]
    lam = rng.standard_normal(12)
    expect = dense.T @ lam
    got = np.zeros(30)
    for u, s in enumerate(uslot):
        got[s] = sum(c[e] * lam[row[e]] for e in range(start[u], start[u + 1]))
    np.testing.assert_allclose(got, expect, atol=1e-14)


@pytest.mark.parametrize("rules", [True, False], ids=["zou-he", "walls"])
def test_ibm_nodes_can_keep_their_pushes(rules):
    """The fused backward step's IBM workgroup keeps the adjoint pushes of the IBM stencil
    nodes and adds their moment rows itself when the transposed row lists are those rows:
    slot (i, x - c_i) lists the density row of node x with coefficient 1 and its momentum rows
    with c_i (nonzero ones), checked here by coordinates with periodic wrap. A region that is
    not the node list or larger than a workgroup, a transposed row list with another
    coefficient, row or order, or a slot list with a slot missing keeps nothing."""
    from lbm.adjoint import keeps_pushes, transpose_rows
    from lbm.ibm import ImmersedForcing

    params = case(dtype="float32", rules=rules)
    ibm = ImmersedForcing(params)
    compiled = ibm.compiled("pull")
    tslot, tstart, trow, tcoeff = transpose_rows(np.asarray(compiled.slots), np.asarray(compiled.coeff))
    ny, nx = params.shape
    nn = ibm.n_nodes
    for q, node in enumerate(ibm.nodes):
        y, x = divmod(int(node), nx)
        for i, (cx, cy) in enumerate(D2Q9.c):
            t = np.searchsorted(tslot, i * nx * ny + ((y - cy) % ny) * nx + (x - cx) % nx)
            rows = list(trow[tstart[t]:tstart[t + 1]])
            coeffs = list(tcoeff[tstart[t]:tstart[t + 1]])
            want = [(a * nn + q, w) for a, w in enumerate((1, cx, cy)) if w != 0]
            assert list(zip(rows, coeffs, strict=True)) == want
    args = (D2Q9, params.shape, ibm.nodes, ibm.nodes, tslot, tstart, trow, tcoeff)
    assert keeps_pushes(*args, 1024)
    assert not keeps_pushes(*args, nn - 1)
    assert not keeps_pushes(D2Q9, params.shape, ibm.nodes[1:], *args[3:], 1024)
    bad = tcoeff.copy()
    bad[7] = 2.0
    assert not keeps_pushes(*args[:7], bad, 1024)
    swapped = trow.copy()
    swapped[[3, 4]] = swapped[[4, 3]]
    assert not keeps_pushes(*args[:6], swapped, tcoeff, 1024)
    assert not keeps_pushes(*args[:4], tslot[1:], tstart[1:] - tstart[1], trow[tstart[1]:], tcoeff[tstart[1]:], 1024)


def _rule_csr(params):
    """Rule targets and the transposed link rules (pull layout) of a case."""
    from lbm.adjoint import transpose_rows
    from lbm.link_rules import compile_rules

    compiled = compile_rules(params.lattice, params.link_rules, "pull", params.solid, params.wall_velocity)
    return (np.asarray(compiled.target),
            *transpose_rows(np.asarray(compiled.values.slots), np.asarray(compiled.values.coeff)))


def test_rule_node_tables_apply_the_transposed_link_rules_bit_for_bit():
    """The fused backward step applies the transposed link rules inside each rule node's
    adjoint step from per-node tables. On random float32 adjoint values, the tables applied
    node by node (a_j = the push of each target direction j, which becomes zero; push i gets
    sum over j of coeff[i][j] a_j in increasing j, added from 0, where the row has entries)
    give the bits of the transposed rules applied to the whole slot array (a_r = lam at each
    target, zeroed, then each term slot += its (rule, coefficient) list in order). The push
    of node x in direction i writes (opp i, x) when x - c_i is solid, else (i, x - c_i),
    computed here by coordinates."""
    from lbm.adjoint import adjoint_rule_nodes

    params = case(dtype="float32", rules=True)
    target, uslot, ustart, urule, ucoeff = _rule_csr(params)
    solid = np.asarray(params.solid, dtype=bool)
    tables = adjoint_rule_nodes(D2Q9, params.shape, solid, target, uslot, ustart, urule, ucoeff)
    assert tables is not None
    nodes, info, coeff = tables
    ny, nx = params.shape
    n = nx * ny
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(9 * n).astype(np.float32)

    expect = lam.copy()
    a = expect[target].copy()
    expect[target] = 0
    for u, s in enumerate(uslot):
        acc = np.float32(0)
        for e in range(ustart[u], ustart[u + 1]):
            acc = np.float32(acc + np.float32(ucoeff[e]) * a[urule[e]])
        expect[s] = np.float32(expect[s] + acc)

    got = lam.copy()
    for k, x in enumerate(nodes):
        y0, x0 = divmod(int(x), nx)
        assert info[k, 0] == x
        slots = []
        for i, (cx, cy) in enumerate(D2Q9.c):
            up = ((y0 - cy) % ny, (x0 - cx) % nx)
            blocked = bool(solid[up])
            assert bool((info[k, 1] >> i) & 1) == blocked
            slots.append(D2Q9.opposite[i] * n + x if blocked else i * n + up[0] * nx + up[1])
        targets = [j for j in range(9) if (info[k, 1] >> (16 + j)) & 1]
        aj = {j: got[slots[j]] for j in targets}
        for j in targets:
            got[slots[j]] = 0
        for i in range(9):
            acc, any_entry = np.float32(0), False
            for j in targets:
                c = np.float32(coeff[k, i, j])
                if c != 0:
                    acc, any_entry = np.float32(acc + c * aj[j]), True
            if any_entry:
                got[slots[i]] = np.float32(got[slots[i]] + acc)
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32))

    # a rule that reads a slot another node writes, or a term slot that is a target, gets no tables
    moved = uslot.copy()
    moved[0] = (moved[0] + nx) % (9 * n)
    assert adjoint_rule_nodes(D2Q9, params.shape, solid, target, moved, ustart, urule, ucoeff) is None
    assert adjoint_rule_nodes(D2Q9, params.shape, solid, target, np.append(uslot[:-1], target[0]), ustart,
                              urule, ucoeff) is None


@pytest.mark.parametrize("rules", [False, True], ids=["walls", "zou-he"])
def test_torch_model_reproduces_the_oracle_drag(rules):
    params = case(rules=rules)
    model = TorchDragModel(params, steps=30, window=30, init_state=oracle_initial_state(params))
    g = torch.zeros((2, *params.shape), dtype=torch.float64)
    x = torch.as_tensor(params.immersed.markers)
    loss = float(model.loss(g, x))
    ref = ReferenceSolver(params)
    ref.initialize(*init_fields(params))
    drags = []
    for _ in range(30):
        ref.step(1)
        drags.append(ref.immersed_force[0])
    assert loss == pytest.approx(np.mean(drags), rel=1e-12)


# ---------------------------------------------------------------- GPU


def make_problem(params, graphs=True, **kw):
    from lbm.adjoint import DragAdjoint

    return DragAdjoint(params, steps=STEPS, window=WINDOW, checkpoint_every=K, init=init_fields(params),
                       graphs=graphs, **kw)


def random_body_field(params, seed=3, scale=1e-6):
    return np.random.default_rng(seed).standard_normal((params.d, *params.shape)) * scale * params.fluid


@pytest.mark.gpu
@pytest.mark.parametrize("rules", [False, True], ids=["walls", "zou-he"])
def test_gpu_forward_drag_matches_the_oracle(rules):
    params = case(rules=rules)
    problem = make_problem(params)
    problem.forward()
    ref = ReferenceSolver(params)
    ref.initialize(*init_fields(params))
    drags = []
    for _ in range(STEPS):
        ref.step(1)
        drags.append(ref.immersed_force[0])
    np.testing.assert_allclose(problem.drag_history, drags, rtol=1e-13, atol=1e-18)


@pytest.mark.gpu
@pytest.mark.parametrize("rules", [False, True], ids=["walls", "zou-he"])
def test_adjoint_matches_float64_torch_autograd(rules):
    params = case(rules=rules)
    g0 = random_body_field(params)
    problem = make_problem(params)
    problem.set_body_force(g0)
    loss = problem.forward()
    grad_g, grad_x = problem.backward()
    model = TorchDragModel(params, STEPS, WINDOW, oracle_initial_state(params))
    g = torch.as_tensor(g0).requires_grad_(True)
    x = torch.as_tensor(params.immersed.markers.copy()).requires_grad_(True)
    ref_loss = model.loss(g, x)
    ref_loss.backward()
    assert loss == pytest.approx(float(ref_loss.detach()), rel=1e-12)
    scale_g = np.abs(g.grad.numpy()).max()
    np.testing.assert_allclose(grad_g, g.grad.numpy(), rtol=0, atol=1e-9 * scale_g)
    np.testing.assert_allclose(grad_x, x.grad.numpy(), rtol=1e-8, atol=1e-9 * np.abs(x.grad.numpy()).max())


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_adjoint_matches_central_finite_differences(dtype):
    """Directional derivatives along a random g-direction and single marker coordinates,
    against float64 central differences of the forward run."""
    params = case(dtype=dtype)
    params64 = case(dtype="float64")
    problem = make_problem(params)
    fd_problem = make_problem(params64)
    g0 = random_body_field(params)
    problem.set_body_force(g0)
    problem.forward()
    grad_g, grad_x = problem.backward()
    direction = random_body_field(params, seed=11, scale=1.0)
    eps = 1e-6

    def loss_at(g, markers):
        fd_problem.set_body_force(g)
        fd_problem.set_markers(markers)
        return fd_problem.forward()

    x0 = params.immersed.markers.copy()
    fd = (loss_at(g0 + eps * direction, x0) - loss_at(g0 - eps * direction, x0)) / (2 * eps)
    adj = float(np.sum(grad_g * direction))
    assert adj == pytest.approx(fd, rel=1e-4)
    for l, a in [(0, 0), (5, 1), (11, 0)]:
        h = 1e-5
        xp, xm = x0.copy(), x0.copy()
        xp[l, a] += h
        xm[l, a] -= h
        fd_x = (loss_at(g0, xp) - loss_at(g0, xm)) / (2 * h)
        assert grad_x[l, a] == pytest.approx(fd_x, rel=1e-4, abs=1e-4 * np.abs(grad_x).max())


@pytest.mark.gpu
@pytest.mark.parametrize("input_dtype", [torch.float32, torch.float64])
def test_drag_loss_autograd_function_gives_the_adjoint_gradients(input_dtype):
    """``drag_loss`` as a torch.autograd.Function on CUDA inputs: the loss equals
    DragAdjoint.forward, ``.grad`` equals DragAdjoint.backward (scaled by the upstream
    gradient), with the inputs' shape (a flattened body field included), dtype and device;
    and one central finite difference of the loss agrees to 1e-4."""
    from lbm.adjoint import drag_loss

    params = case(dtype="float64")
    g0 = random_body_field(params)
    x0 = params.immersed.markers.copy()
    reference = make_problem(params)
    reference.set_body_force(g0)
    ref_loss = reference.forward()
    ref_g, ref_x = reference.backward()

    problem = make_problem(params)
    g = torch.as_tensor(g0, dtype=input_dtype, device="cuda").reshape(params.d, -1).requires_grad_(True)
    x = torch.as_tensor(x0, dtype=input_dtype, device="cuda").requires_grad_(True)
    loss = drag_loss(problem, g, x)
    assert loss.dtype == torch.float64 and loss.device.type == "cuda" and loss.dim() == 0
    (3.0 * loss).backward()
    assert g.grad.shape == g.shape and g.grad.dtype == input_dtype and g.grad.device.type == "cuda"
    assert x.grad.shape == x.shape and x.grad.dtype == input_dtype and x.grad.device.type == "cuda"
    if input_dtype == torch.float64:
        # the same body force and markers: the same forward and adjoint runs, bit for bit
        assert float(loss) == ref_loss
        np.testing.assert_array_equal(g.grad.cpu().numpy(), 3.0 * ref_g.reshape(params.d, -1))
        np.testing.assert_array_equal(x.grad.cpu().numpy(), 3.0 * ref_x)
    else:
        # float32 inputs round g and the markers once on the way in
        assert float(loss) == pytest.approx(ref_loss, rel=1e-5)
        np.testing.assert_allclose(g.grad.cpu().numpy(), (3.0 * ref_g.reshape(params.d, -1)).astype(np.float32),
                                   rtol=1e-4, atol=1e-4 * np.abs(3.0 * ref_g).max())
        np.testing.assert_allclose(x.grad.cpu().numpy(), 3.0 * ref_x, rtol=1e-3, atol=1e-4 * np.abs(3 * ref_x).max())

    eps = 1e-6
    direction = random_body_field(params, seed=11, scale=1.0)
    fd_problem = make_problem(params)

    def loss_at(field):
        with torch.no_grad():
            return float(drag_loss(fd_problem, torch.as_tensor(field, device="cuda"), torch.as_tensor(x0)))

    fd = (loss_at(g0 + eps * direction) - loss_at(g0 - eps * direction)) / (2 * eps)
    assert float(np.sum(ref_g * direction)) == pytest.approx(fd, rel=1e-4)


@pytest.mark.gpu
def test_graph_replay_equals_eager_launches():
    params = case()
    a, b = make_problem(params, graphs=True), make_problem(params, graphs=False)
    g0 = random_body_field(params)
    results = []
    for problem in (a, b):
        problem.set_body_force(g0)
        loss = problem.forward()
        results.append((loss, *problem.backward()))
    # the captured path replays the same kernels: identical bits
    assert results[0][0] == results[1][0]
    np.testing.assert_array_equal(results[0][1], results[1][1])
    np.testing.assert_array_equal(results[0][2], results[1][2])
    # and a second replay of the captured graphs reproduces the first
    a.set_body_force(g0)
    assert a.forward() == results[0][0]
    np.testing.assert_array_equal(a.backward()[0], results[0][1])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_recorded_forcing_equals_recomputed_forcing(dtype):
    """With the per-step IBM fields and intermediates kept (the default) and without them
    (field_store_bytes=0: the segment recompute and the IBM adjoint re-run the forcing),
    the loss and both gradients carry the same bits."""
    params = case(dtype=dtype, rules=True)
    g0 = random_body_field(params)
    results = []
    for store in (16 * 2**30, 0):
        problem = make_problem(params, field_store_bytes=store)
        assert problem.store_fields == bool(store)
        problem.set_body_force(g0)
        results.append((problem.forward(), *problem.backward()))
    assert results[0][0] == results[1][0]
    np.testing.assert_array_equal(results[0][1], results[1][1])
    np.testing.assert_array_equal(results[0][2], results[1][2])


@pytest.mark.gpu
def test_fused_adjoint_step_gives_the_same_bits_with_either_pre_pass():
    """lbm_adj_pull_step_rules with the one-workgroup two-phase pre-pass and with one rule
    per lane of the extra grid layer (rules_direct): the adjoint outputs and the pre-passed
    state agree bit for bit on the Zou-He set, in which no rule reads another's target."""
    from lbm import make_solver
    from lbm.backends._native import FORCE_DENSE, library, pointer

    params = case(dtype="float64", rules=True)
    s = make_solver(params, backend="hip-pull")
    pack = s._rule_passes["pull"]
    assert pack.direct and 0 < pack.n <= 4 * 256
    gen = torch.Generator(device="cuda").manual_seed(11)
    q, n = params.lattice.q, params.n_nodes
    rand = lambda *shape: torch.rand(shape, generator=gen, dtype=torch.float64, device="cuda")  # noqa: E731
    lam_next, f, rf0 = rand(q, n), rand(q, n), rand(q, n)
    field, body = rand(2, n) * 1e-3, rand(2, n) * 1e-6
    wp, wm = params.relaxation_rates
    outs = []
    for direct in (0, 1):
        lam_out = torch.zeros(q, n, dtype=torch.float64, device="cuda")
        grad = torch.zeros(2, n, dtype=torch.float64, device="cuda")
        nu = torch.zeros(2, n, dtype=torch.float64, device="cuda")
        rf = rf0.clone()
        library().call("lbm_adj_pull_step_rules", 0, 1, pointer(lam_next), pointer(lam_out), pointer(f),
                       pointer(s._kind), s._wall_ptr, pointer(field), pointer(body), params.nx, params.ny, 1, wp, wm,
                       0.0, 0.0, 0.0, FORCE_DENSE, 1, pointer(grad), pointer(nu), 64, 4, pointer(rf),
                       pointer(pack.slot), pointer(pack.coeff), pointer(pack.constant), pointer(pack.target), pack.n,
                       pack.m, direct, None)
        torch.cuda.synchronize()
        outs.append((lam_out, grad, nu, rf))
    assert not torch.equal(outs[0][3], rf0)  # the pre-pass wrote its targets
    for a, b in zip(*outs, strict=True):
        assert torch.equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_forward_with_rules_in_the_step_gives_the_same_drag_and_gradients(dtype):
    """The forward run evaluates the link rules inside the pull step (one launch per step
    after the IBM) instead of a pre-pass launch: drag history, loss and gradients are the
    same bits as with the pre-pass and the plain pull."""
    from lbm.adjoint import DragAdjoint

    params = case(dtype=dtype, rules=True)
    g = np.random.default_rng(5).standard_normal((2, *params.shape)) * 1e-6 * params.fluid
    runs = []
    for in_step in (True, False):
        problem = DragAdjoint(params, steps=STEPS, window=WINDOW, checkpoint_every=K, init=init_fields(params))
        assert problem.solver._links is not None
        problem._fused = None  # the two-launch backward step in both runs
        if not in_step:
            problem.solver._links = None
            problem._forcing = None
        problem.set_body_force(g)
        loss = problem.forward()
        runs.append((loss, problem.drag_history, *problem.backward()))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1:], runs[1][1:], strict=True):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("rules", [True, False], ids=["zou-he", "walls"])
def test_fused_backward_step_gives_the_same_bits_as_two_launches(rules):
    """The one-launch backward step (the IBM workgroup runs the adjoint step of its region
    nodes, the forward recompute rides with its link rules inside the pull step and writes
    them back) against the two launches per step: loss and gradients bit for bit, over three
    checkpoint segments."""
    from lbm.adjoint import DragAdjoint

    params = case(dtype="float32", rules=rules)
    g = np.random.default_rng(9).standard_normal((2, *params.shape)) * 1e-6 * params.fluid
    runs = []
    for mode in ("kept", "fused", "two launches"):
        problem = DragAdjoint(params, steps=STEPS, window=WINDOW, checkpoint_every=K, init=init_fields(params))
        assert problem._fused is not None and problem._fused["keep"] == 1
        if mode == "fused":  # the IBM workgroup writes its region's pushes and nu and reads them back
            problem._fused["keep"] = 0
        if mode == "two launches":
            problem._fused = None
        problem.set_body_force(g)
        loss = problem.forward()
        runs.append((loss, *problem.backward()))
    for other in runs[1:]:
        assert runs[0][0] == other[0]
        for a, b in zip(runs[0][1:], other[1:], strict=True):
            assert np.array_equal(a, b)
            assert np.any(a != 0)


@pytest.mark.gpu
@pytest.mark.parametrize("rules", [True, False], ids=["zou-he", "walls"])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_forcing_inside_the_step_gives_the_same_bits_as_two_launches(rules, dtype):
    """The one-launch forward step (the IBM forcing, then the stencil nodes' pull update in
    the forcing's workgroup, every other node in the others) against the forcing launch
    followed by the pull step: drag history, per-step fields and records, loss and gradients
    bit for bit."""
    from lbm.adjoint import DragAdjoint

    params = case(dtype=dtype, rules=rules)
    g = np.random.default_rng(5).standard_normal((2, *params.shape)) * 1e-6 * params.fluid
    runs = []
    for inside in (True, False):
        problem = DragAdjoint(params, steps=STEPS, window=WINDOW, checkpoint_every=K, init=init_fields(params))
        assert problem._forcing is not None
        if not inside:
            problem._forcing = None
        problem.set_body_force(g)
        loss = problem.forward()
        runs.append((loss, problem.drag_history, problem._field_hist.cpu().numpy(), problem._records.cpu().numpy(),
                     *problem.backward()))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1:], runs[1][1:], strict=True):
        assert np.array_equal(a, b)
    assert np.any(runs[0][2] != 0) and np.any(runs[0][3] != 0)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_prefetching_ibm_adjoint_gives_the_same_bits_as_the_generic_kernel(dtype):
    """The IBM adjoint that reads every value it updates ahead of its first barrier (chosen
    when no rule term slot is a rule target or a transposed moment slot) and the generic
    kernel (slots_disjoint = 0) give the same gradients bit for bit, with the link-rule
    adjoint in the same launch, over three checkpoint segments (so both the plain and the
    pull-carrying launch run)."""
    from lbm.adjoint import DragAdjoint

    params = case(dtype=dtype, rules=True)
    g = np.random.default_rng(3).standard_normal((2, *params.shape)) * 1e-6 * params.fluid
    grads = []
    for disjoint in (1, 0):
        problem = DragAdjoint(params, steps=STEPS, window=WINDOW, checkpoint_every=K, init=init_fields(params))
        assert problem._slots_disjoint == 1  # Zou-He rows read neither targets nor IBM slots
        problem._slots_disjoint = disjoint
        problem._fused = None  # the two-launch backward step
        problem.set_body_force(g)
        problem.forward()
        grads.append(problem.backward())
    for a, b in zip(*grads, strict=True):
        assert np.array_equal(a, b)
        assert np.any(a != 0)


@pytest.mark.gpu
def test_full_size_ibm_channel_gradients_match_central_finite_differences():
    """The differentiable channel at its full 512 x 128 size (Zou-He inlet and outlet, walls,
    IBM cylinder R = 12 with 3 multi-direct passes, Re 100) in float64, over 400 steps in
    four checkpoint segments of 100 (so the pipelined recompute and its riding pre-pass
    run): the directional derivative along a random body-force direction and two marker
    coordinates agree with central finite differences of the forward run to 1e-4."""
    from lbm.adjoint import DragAdjoint
    from lbm.cases.ibm_channel import ibm_channel_params, initial_fields

    params = ibm_channel_params(dtype="float64")
    assert params.shape == (128, 512) and len(params.link_rules) > 0

    def problem():
        return DragAdjoint(params, steps=400, window=200, checkpoint_every=100, init=initial_fields(params))

    adj_problem, fd_problem = problem(), problem()
    g0 = np.random.default_rng(0).standard_normal((2, *params.shape)) * 1e-7 * params.fluid
    adj_problem.set_body_force(g0)
    loss = adj_problem.forward()
    grad_g, grad_x = adj_problem.backward()
    assert np.isfinite(loss) and np.all(np.isfinite(grad_g)) and np.all(np.isfinite(grad_x))

    def loss_at(g, markers):
        fd_problem.set_body_force(g)
        fd_problem.set_markers(markers)
        return fd_problem.forward()

    x0 = params.immersed.markers.copy()
    assert loss_at(g0, x0) == loss  # the same forward, bit for bit
    direction = np.random.default_rng(11).standard_normal((2, *params.shape)) * params.fluid
    eps = 1e-6
    fd = (loss_at(g0 + eps * direction, x0) - loss_at(g0 - eps * direction, x0)) / (2 * eps)
    assert float(np.sum(grad_g * direction)) == pytest.approx(fd, rel=1e-4)
    for l, a in [(0, 0), (37, 1)]:
        h = 1e-5
        xp, xm = x0.copy(), x0.copy()
        xp[l, a] += h
        xm[l, a] -= h
        fd_x = (loss_at(g0, xp) - loss_at(g0, xm)) / (2 * h)
        assert grad_x[l, a] == pytest.approx(fd_x, rel=1e-4, abs=1e-4 * np.abs(grad_x).max())


@pytest.mark.gpu
def test_config_size_float32_gradients_match_float64_finite_differences():
    """The differentiable configuration as stated: D2Q9 float32 channel 512 x 128, IBM
    cylinder R = 12 (3 multi-direct passes), Re 100, 4,000 steps, loss = mean drag over the
    last 2,000, checkpoints every 100, body force N(0, 1e-7). The float32 adjoint's
    directional derivative along a random body-force direction, and one marker coordinate,
    agree with float64 central finite differences of the float64 forward run to 1e-4
    relative."""
    from lbm.adjoint import DragAdjoint
    from lbm.cases.ibm_channel import ibm_channel_params, initial_fields

    def problem(dtype):
        params = ibm_channel_params(dtype=dtype)
        return params, DragAdjoint(params, steps=4000, window=2000, checkpoint_every=100,
                                   init=initial_fields(params))

    params, adj = problem("float32")
    params64, fd_problem = problem("float64")
    g0 = np.random.default_rng(0).standard_normal((2, *params.shape)) * 1e-7 * params.fluid
    adj.set_body_force(g0)
    loss32 = adj.forward()
    grad_g, grad_x = adj.backward()
    assert np.all(np.isfinite(grad_g)) and np.all(np.isfinite(grad_x))

    def loss_at(g, markers):
        fd_problem.set_body_force(g)
        fd_problem.set_markers(markers)
        return fd_problem.forward()

    x0 = params64.immersed.markers.copy()
    assert loss_at(g0, x0) == pytest.approx(loss32, rel=1e-4)
    direction = np.random.default_rng(11).standard_normal((2, *params.shape)) * params.fluid
    eps = 1e-6
    fd = (loss_at(g0 + eps * direction, x0) - loss_at(g0 - eps * direction, x0)) / (2 * eps)
    assert float(np.sum(grad_g * direction)) == pytest.approx(fd, rel=1e-4)
    l, a = 0, 0
    h = 1e-5
    xp, xm = x0.copy(), x0.copy()
    xp[l, a] += h
    xm[l, a] -= h
    fd_x = (loss_at(g0, xp) - loss_at(g0, xm)) / (2 * h)
    assert grad_x[l, a] == pytest.approx(fd_x, rel=1e-4, abs=1e-4 * np.abs(grad_x).max())
