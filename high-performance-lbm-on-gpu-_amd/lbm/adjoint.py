"""Gradients of the time-averaged drag of an immersed body by a hand-written discrete adjoint.

No counterpart in the reference (which has no autograd or checkpointing). The forward
model is the HIP pull step of ``HipSolver`` with a dense per-node body force g(x) added
to the uniform force:

    per step t:  h^t, D^t = IBM(f^t; X)              (D^t: x component of the body force)
                 f_s = stream + bounce-back + rules(f^t)
                 f^{t+1} = collide(f_s, F0 + g + h^t)  at fluid nodes
    loss L = mean of D^t over the last ``window`` steps.

``DragAdjoint.backward`` returns dL/dg (d, *shape) and dL/dX (n_markers, d), both float64.
The adjoint runs backwards over segments of ``checkpoint_every`` steps: each segment is
recomputed forward from its checkpoint (states kept in HBM), then swept backwards with
``lbm_adj_pull_step`` (transposed collision Jacobian + push) and ``lbm_ibm_adjoint``
(transposed link rules and reverse multi-direct forcing in one launch). While HBM allows
(``field_store_bytes``), the forward run keeps each step's IBM field and the intermediates
of its forcing iterations, so neither the segment recompute nor the IBM adjoint re-runs the
forcing. dL/dX follows from the accumulated dL/dw and the derivative of the Roma kernel.

Segments are captured once into HIP graphs (through ``torch.cuda.CUDAGraph``) and replayed,
so the Python loop does not set the step rate. The forward recompute of segment s - 1 runs
inside the launches of segment s's backward sweep (``lbm_adj_pull_step_rules``,
``lbm_ibm_adjoint_pull``): both are small, latency-bound launches, and pairing them halves
the launches per backward step. ``DragLoss`` wraps the whole thing as a
``torch.autograd.Function``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from lbm.backends import graph_capture
from lbm.backends._native import FORCE_DENSE, pointer
from lbm.backends.hip_solver import KIND_REGION, HipSolver
from lbm.geometry import neighbour_index
from lbm.ibm import STENCIL, roma, roma_derivative
from lbm.link_rules import compile_rules


def transpose_rows(slots: np.ndarray, coeff: np.ndarray):
    """CSR of the transposed affine rows: distinct slots, start offsets, (row, coeff) entries."""
    rows, terms = np.nonzero(coeff)
    s = slots[rows, terms]
    order = np.lexsort((terms, rows, s))
    s, rows, c = s[order], rows[order], coeff[rows[order], terms[order]]
    unique, start = np.unique(s, return_index=True)
    start = np.append(start, len(s)).astype(np.int32)
    return unique.astype(np.int64), start, rows.astype(np.int32), c


def keeps_pushes(lattice, shape, region, nodes, tslot, tstart, trow, tcoeff, lanes: int) -> bool:
    """Whether the fused backward step's IBM workgroup can keep its region's adjoint pushes:
    the region is the IBM node list itself, at most one node per lane, and the transposed
    slots are exactly the pushes of those nodes, the push of node x in direction i writing
    slot (i, x - c_i) (flat i * n + (x - c_i), periodic wrap), whose transposed row list
    holds the moment rows of node x only: rows a * nn + x (nn IBM nodes; a = 0 density with
    coefficient 1, a = 1 .. d momentum with c_i,a-1), the nonzero ones, in increasing a."""
    region, nodes, tslot = np.asarray(region), np.asarray(nodes), np.asarray(tslot)
    nn, q = len(nodes), lattice.q
    if not np.array_equal(region, nodes) or nn > lanes or len(tslot) != q * nn:
        return False
    c = np.asarray(lattice.c)
    weight = np.vstack([np.ones(q), c.T])  # (d + 1, q): coefficient of direction i in row block a
    upstream = np.stack([neighbour_index(region, tuple(-c[i]), shape) for i in range(q)], axis=1)
    slots = (np.arange(q) * int(np.prod(shape)) + upstream).ravel()
    at = np.minimum(np.searchsorted(tslot, slots), len(tslot) - 1)  # tslot is sorted and distinct
    if not (np.all(np.diff(tslot) > 0) and np.array_equal(tslot[at], slots)):
        return False
    tstart, trow, tcoeff = np.asarray(tstart), np.asarray(trow), np.asarray(tcoeff)
    for x in range(nn):
        for i in range(q):
            t = at[x * q + i]
            blocks = np.flatnonzero(weight[:, i])
            if not (np.array_equal(trow[tstart[t]:tstart[t + 1]], blocks * nn + x)
                    and np.array_equal(tcoeff[tstart[t]:tstart[t + 1]], weight[blocks, i])):
                return False
    return True


def slot_writers(lattice, shape, solid, slots):
    """The adjoint push that writes each lam slot (k, y) of a fluid node y: (node, direction)
    = (y + c_k, k) when y + c_k is fluid, else y's own bounce-back (y, opp k)."""
    n = int(np.prod(shape))
    k, y = np.asarray(slots) // n, np.asarray(slots) % n
    downstream = neighbour_index(y, tuple(np.asarray(lattice.c)[k].T), shape)
    fluid = ~solid[downstream]
    return np.where(fluid, downstream, y), np.where(fluid, k, np.asarray(lattice.opposite)[k])


def adjoint_rule_nodes(lattice, shape, solid, target, uslot, ustart, urule, ucoeff):
    """The link-rule node tables of the fused backward step, or None when the rule set does
    not allow them: every rule's target and term slots written by the adjoint step of one
    node (the rule's node), no term slot a target, and each transposed row's entries in
    increasing target direction with no (slot, rule) pair twice. Returns (nodes, info,
    coeff): the rule nodes (sorted), info (count, 4) uint32 = {node, blocked directions |
    target directions << 16, 0, 0} (blocked: x - c_i solid), coeff (count, q, q) with
    coeff[k, i, j] = the coefficient of the rule whose target push is j in the transposed row
    of the slot push i writes."""
    solid = np.asarray(solid, dtype=bool).ravel()
    target, uslot = np.asarray(target), np.asarray(uslot)
    ustart, urule, ucoeff = np.asarray(ustart), np.asarray(urule), np.asarray(ucoeff)
    if len(target) == 0 or np.intersect1d(uslot, target).size:
        return None
    tx, tj = slot_writers(lattice, shape, solid, target)  # per rule
    ux, ui = slot_writers(lattice, shape, solid, uslot)  # per term slot
    entry_slot = np.repeat(np.arange(len(uslot)), np.diff(ustart))
    if np.any(tx[urule] != ux[entry_slot]) or np.any(solid[tx]):
        return None
    same_slot = entry_slot[1:] == entry_slot[:-1]
    if np.any(same_slot & (tj[urule][1:] <= tj[urule][:-1])):  # increasing j, no rule twice in a row
        return None
    nodes, node_of_rule = np.unique(tx, return_inverse=True)
    q = lattice.q
    coeff = np.zeros((len(nodes), q, q))
    coeff[node_of_rule[urule], ui[entry_slot], tj[urule]] = ucoeff
    targets = np.zeros(len(nodes), dtype=np.int64)
    np.bitwise_or.at(targets, node_of_rule, np.left_shift(1, tj))
    if np.any(np.bincount(node_of_rule, minlength=len(nodes)) != [bin(t).count("1") for t in targets]):
        return None  # two rules of one node with the same target push
    c = np.asarray(lattice.c)
    blocked = np.zeros(len(nodes), dtype=np.int64)
    for i in range(q):
        blocked |= solid[neighbour_index(nodes, tuple(-c[i]), shape)].astype(np.int64) << i
    info = np.zeros((len(nodes), 4), dtype=np.uint32)
    info[:, 0] = nodes
    info[:, 1] = blocked | (targets << 16)
    return nodes, info, coeff


def marker_gradient(markers: np.ndarray, stencil_base: np.ndarray, grad_w: np.ndarray) -> np.ndarray:
    """dL/dX from dL/dw, w_ls = prod_a phi(p_ls,a - X_l,a): dw/dX_l,a = -phi'(r_a) prod_{b!=a} phi(r_b)."""
    n, d = markers.shape
    offsets = np.stack(np.meshgrid(*(np.arange(STENCIL),) * d, indexing="ij"), axis=-1).reshape(-1, d)
    r = (stencil_base[:, None, :] + offsets[None, :, :]) - markers[:, None, :]
    phi = roma(r)
    dphi = roma_derivative(r)
    grad = np.zeros((n, d))
    for a in range(d):
        others = np.prod(np.delete(phi, a, axis=-1), axis=-1)
        grad[:, a] = np.sum(grad_w * (-dphi[..., a] * others), axis=1)
    return grad


class DragAdjoint:
    """Checkpointed forward/adjoint of the mean IBM drag (see module docstring).

    Args:
        params: ``SimulationParams`` with an immersed body (drag = its x body force).
        steps: time steps of the forward run.
        window: the loss averages the drag of steps ``steps - window`` .. ``steps - 1``.
        checkpoint_every: segment length k (states of one segment are recomputed).
        init: initial macroscopic fields (rho, ux, uy[, uz]); default rest.
        graphs: capture segments into HIP graphs (needs ``steps % k == 0`` and even k).
        field_store_bytes: HBM budget for the per-step IBM fields of the forward run.
    """

    def __init__(self, params, steps: int, window: int, checkpoint_every: int = 100, init=None,
                 block=None, graphs: bool = True, field_store_bytes: int = 16 * 2**30):
        if params.immersed is None:
            raise ValueError("the drag functional needs an immersed body (params.immersed)")
        if not 0 < window <= steps:
            raise ValueError("need 0 < window <= steps")
        k = int(checkpoint_every)
        if k <= 0:
            raise ValueError("checkpoint_every must be positive")
        solver = HipSolver(params, variant="pull", block=block)
        self.solver = s = solver
        self.params = p = params
        self.steps, self.window, self.k = steps, window, k
        if k % 2:
            raise ValueError("checkpoint_every must be even (the adjoint buffers alternate per step)")
        self.graphs = graphs and steps % k == 0
        torch = s._torch
        self._torch = torch
        dev, tdt = s.device, s._torch_dtype
        d, q, n = p.d, p.lattice.q, s._n
        self.init = init if init is not None else (1.0,)
        s.set_body_force_field(np.zeros((d, *p.shape)))
        s._force_mode = FORCE_DENSE
        self.n_segments = -(-steps // k)
        self._ckpt = torch.zeros((self.n_segments, q, n), dtype=tdt, device=dev)
        # Two segment buffers: one segment is swept backwards while the previous one is
        # recomputed forward inside the same launches.
        self._segs = (torch.zeros((k + 1, q, n), dtype=tdt, device=dev),
                      torch.zeros((k + 1, q, n), dtype=tdt, device=dev))
        # Each forward step records its IBM force field into its own pre-zeroed dense buffer
        # (the IBM kernel writes only the stencil nodes), so the segment recompute reuses it
        # instead of re-running the IBM; above ``field_store_bytes`` the recompute re-runs it.
        # With the fields, each step also keeps the IBM intermediates (lbm_ibm_record_size
        # bytes), which the IBM adjoint then reads instead of recomputing the iterations.
        ibm = s._ibm
        rec = ctypes.c_int64()
        s._lib.call("lbm_ibm_record_size", s._lattice_code, s._dtype_code, ibm.n_nodes, ibm.n_markers,
                    ibm.iterations, ctypes.byref(rec))
        rec_elems = rec.value // s._np_dtype.itemsize
        field_bytes = steps * (d * n + rec_elems) * s._np_dtype.itemsize
        self.store_fields = field_bytes <= field_store_bytes
        self._field_hist = torch.zeros((steps if self.store_fields else 2 * k, d, n), dtype=tdt, device=dev)
        self._records = torch.zeros((steps if self.store_fields else 0, rec_elems), dtype=tdt, device=dev)
        self._start = torch.zeros((q, n), dtype=tdt, device=dev)
        self._drag = torch.zeros((steps, d), dtype=tdt, device=dev)
        self._drag_stage = torch.zeros((k, d), dtype=tdt, device=dev)
        self._bf_scratch = torch.zeros(d, dtype=tdt, device=dev)
        weights = np.zeros((steps, d))
        weights[steps - window:, 0] = 1.0 / window
        self._body_weight = torch.as_tensor(weights, device=dev)
        self._lam = torch.zeros((2, q, n), dtype=tdt, device=dev)
        self._grad_body = torch.zeros((d, n), dtype=torch.float64, device=dev)
        self._nu = torch.zeros((d, n), dtype=tdt, device=dev)
        self._setup_ibm_adjoint()
        self._setup_rule_adjoint()
        self._setup_fused_step()
        self._setup_forcing_in_step()
        self._graphs = {}
        self.set_initial_state(*self.init)

    # ------------------------------------------------------------------ set-up

    def _setup_ibm_adjoint(self) -> None:
        s, torch, dev = self.solver, self._torch, self.solver.device
        ibm = s._ibm
        mp = s._ibm_pack("pull")
        compiled = ibm.compiled("pull")
        tslot, tstart, trow, tcoeff = transpose_rows(np.asarray(compiled.slots), np.asarray(compiled.coeff))
        if s.params.link_rules is not None and len(s.params.link_rules):
            targets = compile_rules(s.params.lattice, s.params.link_rules, "pull", s.params.solid,
                                    s.params.wall_velocity).target
            if np.intersect1d(targets, tslot).size:
                raise NotImplementedError("IBM moment rows read a link-rule target slot; move the body away")
        self._mp = mp
        self._tslot = torch.as_tensor(tslot, device=dev)
        self._tstart = torch.as_tensor(tstart, device=dev)
        self._trow = torch.as_tensor(trow, device=dev)
        self._tcoeff = torch.as_tensor(tcoeff.astype(s._np_dtype), device=dev)
        self._grad_w = torch.zeros((ibm.n_markers, STENCIL**s.params.d), dtype=torch.float64, device=dev)
        size = ctypes.c_int64()
        s._lib.call("lbm_ibm_adjoint_scratch_size", s._lattice_code, s._dtype_code, ibm.n_nodes, ibm.n_markers,
                    ibm.iterations, ctypes.byref(size))
        self._ibm_adj_scratch = torch.zeros(size.value // s._np_dtype.itemsize + 1, dtype=s._torch_dtype, device=dev)

    def _setup_rule_adjoint(self) -> None:
        s, torch, dev = self.solver, self._torch, self.solver.device
        pack = s._rule_passes.get("pull")
        self._rules = None
        self._slots_disjoint = 1
        if pack is None:
            return
        compiled = pack.source
        uslot, ustart, urule, ucoeff = transpose_rows(np.asarray(compiled.slots), np.asarray(compiled.coeff))
        # Rule term slots that are neither rule targets nor transposed IBM slots let the IBM
        # adjoint read every value it updates ahead of its first barrier (slots_disjoint).
        targets = pack.target.cpu().numpy()
        self._slots_disjoint = int(not (np.intersect1d(uslot, targets).size
                                        or np.intersect1d(uslot, self._tslot.cpu().numpy()).size))
        self._rules = {
            "target": pack.target,
            "n": pack.n,
            "uslot": torch.as_tensor(uslot, device=dev),
            "ustart": torch.as_tensor(ustart, device=dev),
            "urule": torch.as_tensor(urule, device=dev),
            "ucoeff": torch.as_tensor(ucoeff.astype(s._np_dtype), device=dev),
            "nu": len(uslot),
            "scratch": torch.zeros(max(1, pack.n), dtype=s._torch_dtype, device=dev),
        }

    def _setup_fused_step(self) -> None:
        """The one-launch backward step (``lbm_adj_fused_step``) when the library takes this
        problem, the forward recompute can run its link rules inside the pull step (a rule set
        in which no rule reads another's target) and the rule adjoint can run inside the rule
        nodes' own adjoint step (``adjoint_rule_nodes``). Its IBM region: the writers of the
        slots the IBM adjoint reads or updates (the stencil nodes), apart from the rule nodes.
        When the region is the IBM node list itself, the IBM workgroup keeps its region's pushes
        and nu (``keeps_pushes``)."""
        s, torch = self.solver, self._torch
        self._fused = None
        ibm, r = s._ibm, self._rules
        pack = s._rule_passes.get("pull")
        if pack is not None and (s._links is None or not pack.direct or not self._slots_disjoint):
            return
        bx, by = s.block
        ok = ctypes.c_int()
        s._lib.call("lbm_adj_fused_supported", s._lattice_code, s._dtype_code, ibm.n_nodes, ibm.n_markers,
                    ibm.iterations, len(self._tslot), bx, by, ctypes.byref(ok))
        if not ok.value:
            return
        p, n = s.params, s._n
        solid = np.zeros(n, dtype=bool) if p.solid is None else np.asarray(p.solid, dtype=bool).ravel()
        region = np.unique(slot_writers(p.lattice, p.shape, solid, self._tslot.cpu().numpy())[0])
        rule_nodes = np.zeros(0, dtype=np.int64)
        info = coeff = None
        if r is not None:
            tables = adjoint_rule_nodes(p.lattice, p.shape, solid, r["target"].cpu().numpy(),
                                        r["uslot"].cpu().numpy(), r["ustart"].cpu().numpy(),
                                        r["urule"].cpu().numpy(), r["ucoeff"].cpu().numpy())
            if tables is None:
                return
            rule_nodes, info, coeff = tables
        both = np.concatenate([region, rule_nodes])
        if solid[both].any() or np.intersect1d(region, rule_nodes).size:
            return
        if (s._kind.cpu().numpy()[region] & 4).any():  # the IBM region takes no near-solid node
            return
        kind = s._kind.clone()
        kind[torch.as_tensor(both, device=s.device)] |= KIND_REGION
        tslot = self._tslot.cpu().numpy()
        keep = keeps_pushes(p.lattice, p.shape, region, ibm.nodes, tslot, self._tstart.cpu().numpy(),
                            self._trow.cpu().numpy(), self._tcoeff.cpu().numpy(), bx * by)
        self._fused = {"region": torch.as_tensor(region.astype(np.int32), device=s.device), "kind": kind,
                       "keep": int(keep), "rule_nodes": len(rule_nodes),
                       "rule_info": None if info is None else torch.as_tensor(info.view(np.int32), device=s.device),
                       "rule_coeff": None if coeff is None else torch.as_tensor(coeff.astype(s._np_dtype),
                                                                                device=s.device)}

    def _setup_forcing_in_step(self) -> None:
        """The one-launch forward step (``lbm_ibm_pull_step``: the IBM forcing and the pull step
        with its link rules) when the solver can take it with the forcing's record
        (``HipSolver._forcing_in_step``)."""
        s = self.solver
        self._forcing = s._forcing_in_step(record=True) if s._forcing_in_step() is not None else None

    # ------------------------------------------------------------------ controls

    def set_initial_state(self, *fields) -> None:
        self.solver.initialize(*fields)
        self._start.copy_(self.solver._f)

    def set_body_force(self, field) -> None:
        """Dense body force g, shape (d, *shape) (NumPy or torch, any device)."""
        t = self._torch.as_tensor(field, dtype=self.solver._torch_dtype)
        self.solver._body.copy_(t.reshape(self.solver._body.shape))

    def set_markers(self, markers) -> None:
        """New marker positions (same count); the stencil node set must stay the same."""
        s = self.solver
        body = s.params.immersed.with_markers(np.asarray(markers, dtype=np.float64))
        old = s._ibm
        stencil, weights = body.stencil(s.params.shape)
        nodes, inverse = np.unique(stencil, return_inverse=True)
        if not (np.array_equal(nodes, old.nodes) and np.array_equal(inverse.reshape(stencil.shape), old.inverse)):
            raise ValueError("the markers moved across a grid line: the stencil changed")
        old.body = body
        old.weights = weights
        s._ibm_weights.copy_(self._torch.as_tensor(weights.astype(s._np_dtype)))

    @property
    def markers(self) -> np.ndarray:
        return np.array(self.solver._ibm.body.markers)

    # ------------------------------------------------------------------ launches

    def _ibm_forward(self, f, field, body_out, record=None) -> None:
        s, ibm, mp = self.solver, self.solver._ibm, self._mp
        fx, fy, fz = s._force3
        s._call("lbm_ibm_force", s._lattice_code, s._dtype_code, pointer(f), pointer(mp.slot), pointer(mp.coeff),
                pointer(mp.constant), mp.m, pointer(s._ibm_nodes), ibm.n_nodes, pointer(s._ibm_inverse),
                pointer(s._ibm_weights), pointer(s._ibm_areas), ibm.n_markers, pointer(s._ibm_start),
                pointer(s._ibm_entry), ibm.iterations, fx, fy, fz, s._n, pointer(field), body_out,
                pointer(s._ibm_scratch), pointer(record), s._stream)

    def _rules_forward(self, f) -> None:
        s = self.solver
        pack = s._rule_passes.get("pull")
        if pack is not None:
            name, args = s._rules_call(pack, pointer(f), s._stream)
            s._call(name, *args)

    def _pull(self, src, dst, field) -> None:
        s = self.solver
        s._call("lbm_pull_step", s._lattice_code, s._dtype_code, pointer(src), pointer(dst), pointer(s._kind),
                s._wall_ptr, pointer(field), pointer(s._body), *s._step_args())

    def _forced_step(self, src, dst, field, body_out, record=None) -> None:
        """The IBM forcing of ``src`` and the pull step to ``dst`` in one launch: the same
        field, drag, record and next state as ``_ibm_forward`` followed by ``_step_forward``."""
        s = self.solver
        name, args = s._ibm_step_call(pointer(src), pointer(dst), self._forcing, s._stream, field=pointer(field),
                                      body_force=body_out, record=pointer(record))
        s._call(name, *args)

    def _step_forward(self, src, dst, field) -> None:
        """One forward pull step from ``src`` to ``dst``. With link rules it evaluates them
        inside the step (``lbm_pull_step_links``): ``dst`` gets the same bits as the rule
        pre-pass followed by the pull, and only ``src`` keeps its rule targets unwritten. The
        forward run never reads those slots again (the IBM moment rows read none, and the next
        step overwrites the buffer), so drag history, records and checkpoints are the same.
        The backward sweep's recompute keeps the pre-pass: the adjoint step reads the rule
        values from the states it recomputes."""
        s = self.solver
        if s._links is not None and s.block[0] * s.block[1] >= 64:
            name, args = s._pull_links_call(pointer(src), pointer(dst), s._stream, field=pointer(field))
            s._call(name, *args)
            return
        self._rules_forward(src)
        self._pull(src, dst, field)

    def _forward_core(self, length: int, drag, fields, records=None) -> None:
        """``length`` steps from solver._f (ending there: length is even or a copy closes it).

        ``fields`` is the per-step field buffer (j-th step writes fields[j]) or None to use
        the solver's single field buffer; ``records[j]`` (optional) receives the j-th step's
        IBM intermediates."""
        s = self.solver
        a, b = s._f, s._f_next
        for j in range(length):
            cur, nxt = (a, b) if j % 2 == 0 else (b, a)
            field = s._field if fields is None else fields[j]
            record = None if records is None else records[j]
            if self._forcing is not None and (s._links is not None or "pull" not in s._rule_passes):
                self._forced_step(cur, nxt, field, drag[j].data_ptr(), record)
                continue
            self._ibm_forward(cur, field, drag[j].data_ptr(), record)
            self._step_forward(cur, nxt, field)
        if length % 2:
            a.copy_(b)

    def _recompute(self, buf, fields, j0: int, j1: int) -> None:
        """Forward steps j0 .. j1 - 1 of a segment held in ``buf`` (buf[j] -> buf[j + 1]).

        ``fields[j]`` is the IBM field of step j; recorded fields spare the IBM."""
        for j in range(j0, j1):
            if not self.store_fields:
                self._ibm_forward(buf[j], fields[j], self._bf_scratch.data_ptr())
            self._rules_forward(buf[j])
            self._pull(buf[j], buf[j + 1], fields[j])

    def _sweep(self, buf, fields, records, weights, length: int, partner) -> None:
        """The backward sweep of a recomputed segment (states buf[0 .. length]).

        ``partner`` = (buffer, fields, length) of the previous segment, or None. Its forward
        recompute rides in the same launches: partner step jp = length - 1 - j goes with
        backward step j, its link-rule pre-pass in the adjoint step's launch and its pull in
        the IBM adjoint's launch; partner steps beyond ``length`` run after the sweep. The
        adjoint buffers alternate with the parity of the step inside the segment; segments
        start at even steps (k is even), so the parity is the global one.
        """
        s = self.solver
        L, D = s._lattice_code, s._dtype_code
        ibm, mp = s._ibm, self._mp
        fx, fy, fz = s._force3
        wp, wm = s._rates
        bx, by = self.adjoint_block
        nx, ny, nz = s.params.nx, s.params.ny, s._nz
        pack = s._rule_passes.get("pull")
        # the recompute's pre-pass rides in the adjoint step's launch: one rule per lane of an
        # extra grid layer at the adjoint's own block shape when no rule reads another's
        # target, else one workgroup of up to 4 rules per lane
        layer = -(-nx // bx) * -(-ny // by) * bx * by
        direct = pack is not None and pack.direct and pack.n <= layer
        fbx, fby = (bx, by) if direct else self.fused_block
        fuse_rules = pack is not None and (direct or pack.n <= 4 * fbx * fby)
        r = self._rules or {"target": None, "n": 0, "uslot": None, "ustart": None, "urule": None,
                            "ucoeff": None, "nu": 0, "scratch": None}
        for j in reversed(range(length)):
            lam_next, lam_out = self._lam[(j + 1) % 2], self._lam[j % 2]
            jp = length - 1 - j
            with_partner = partner is not None and jp < partner[2]
            if with_partner:
                buf_p, fields_p, _ = partner
                if not self.store_fields:
                    self._ibm_forward(buf_p[jp], fields_p[jp], self._bf_scratch.data_ptr())
            if self._fused is not None:
                following = None
                if j > 0 and records is not None:
                    following = (buf[j - 1], records[j - 1], fields[j - 1])
                self._fused_step(buf[j], fields[j], None if records is None else records[j], weights[j], lam_next,
                                 lam_out, (partner[0][jp], partner[0][jp + 1], partner[1][jp]) if with_partner else None,
                                 following)
                continue
            adj = (L, D, pointer(lam_next), pointer(lam_out), pointer(buf[j]), pointer(s._kind), s._wall_ptr,
                   pointer(fields[j]), pointer(s._body), nx, ny, nz, wp, wm, fx, fy, fz, FORCE_DENSE, s._mask_mode,
                   pointer(self._grad_body), pointer(self._nu), bx, by)
            if with_partner and fuse_rules:
                s._call("lbm_adj_pull_step_rules", *adj[:-2], fbx, fby, pointer(buf_p[jp]), pointer(pack.slot),
                        pointer(pack.coeff), pointer(pack.constant), pointer(pack.target), pack.n, pack.m,
                        int(direct), s._stream)
            else:
                if with_partner:
                    self._rules_forward(buf_p[jp])
                s._call("lbm_adj_pull_step", *adj, s._stream)
            ibm_args = (L, D, pointer(buf[j]), pointer(mp.slot), pointer(mp.coeff), pointer(mp.constant), mp.m,
                        pointer(s._ibm_nodes), ibm.n_nodes, pointer(s._ibm_inverse), pointer(s._ibm_weights),
                        pointer(s._ibm_areas), ibm.n_markers, pointer(s._ibm_start), pointer(s._ibm_entry),
                        ibm.iterations, fx, fy, fz, s._n, pointer(self._nu), weights[j].data_ptr(),
                        pointer(lam_out), pointer(self._tslot), pointer(self._tstart), pointer(self._trow),
                        pointer(self._tcoeff), len(self._tslot), pointer(self._grad_w), pointer(self._ibm_adj_scratch),
                        pointer(r["target"]), r["n"], pointer(r["uslot"]), pointer(r["ustart"]), pointer(r["urule"]),
                        pointer(r["ucoeff"]), r["nu"], pointer(r["scratch"]), self._slots_disjoint,
                        None if records is None else pointer(records[j]))
            if with_partner:
                s._call("lbm_ibm_adjoint_pull", *ibm_args, pointer(buf_p[jp]), pointer(buf_p[jp + 1]),
                        pointer(s._kind), s._wall_ptr, pointer(fields_p[jp]), pointer(s._body), nx, ny, nz,
                        wp, wm, FORCE_DENSE, s._mask_mode, s._stream)
            else:
                s._call("lbm_ibm_adjoint", *ibm_args, s._stream)
        if partner is not None and partner[2] > length:
            self._recompute(partner[0], partner[1], length, partner[2])

    def _fused_step(self, f, field, record, weight, lam_next, lam_out, partner, following=None) -> None:
        """One backward step in one launch; ``partner`` = (source, destination, field) of the
        forward recompute step riding along, or None; ``following`` = (state, record, field) of
        the backward step after this one (read ahead into the cache), or None."""
        s = self.solver
        ibm, mp, r, fz = s._ibm, self._mp, self._rules, self._fused
        fx, fy, fz_ = s._force3
        wp, wm = s._rates
        bx, by = s.block
        if partner is not None and r is not None and s._links is None:
            raise RuntimeError("the fused backward step needs the solver's in-step link rules for the recompute")
        if partner is None or s._links is None:
            links = (None, None, 0, None, None, None, 0, 1)
        else:
            lk = s._links
            links = (pointer(lk.node_info), pointer(lk.node_first), lk.nodes, pointer(lk.slot), pointer(lk.coeff),
                     pointer(lk.constant), lk.n, lk.m)
        src, dst, pfield = (None, None, None) if partner is None else tuple(pointer(t) for t in partner)
        s._call("lbm_adj_fused_step", s._lattice_code, s._dtype_code, pointer(f), pointer(mp.slot), pointer(mp.coeff),
                pointer(mp.constant), mp.m, pointer(s._ibm_nodes), ibm.n_nodes, pointer(s._ibm_inverse),
                pointer(s._ibm_weights), pointer(s._ibm_areas), ibm.n_markers, pointer(s._ibm_start),
                pointer(s._ibm_entry), ibm.iterations, fx, fy, fz_, s._n, pointer(self._nu), weight.data_ptr(),
                pointer(lam_out), pointer(self._tslot), pointer(self._tstart), pointer(self._trow),
                pointer(self._tcoeff), len(self._tslot), pointer(self._grad_w),
                None if record is None else pointer(record), pointer(lam_next), pointer(fz["kind"]), s._wall_ptr,
                pointer(field), pointer(s._body), s.params.nx, s.params.ny, s._nz, wp, wm, pointer(self._grad_body),
                pointer(fz["region"]), len(fz["region"]), fz["keep"], pointer(fz["rule_info"]),
                pointer(fz["rule_coeff"]), fz["rule_nodes"], bx, by, src, dst, pfield, *links,
                *((None, None, None) if following is None else tuple(pointer(t) for t in following)), s._stream)

    @property
    def adjoint_block(self) -> tuple[int, int]:
        bx = min(256, 64 * -(-self.params.nx // 64))
        return bx, max(1, min(self.params.ny, 256 // bx))

    @property
    def fused_block(self) -> tuple[int, int]:
        """Block of the adjoint step that carries a link-rule pre-pass: up to 1024 lanes for
        D2Q9 float32 (one lane per rule), 512 for D2Q9 float64, 256 for D3Q19."""
        bx, _ = self.adjoint_block
        lanes = 256 if self.params.d == 3 else (1024 if self.solver._np_dtype.itemsize == 4 else 512)
        return bx, max(1, min(self.params.ny, lanes // bx))

    # ------------------------------------------------------------------ graphs

    def _replay(self, key, launch) -> None:
        """Run ``launch`` eagerly the first time, then capture it and replay the graph."""
        if not self.graphs:
            launch()
            return
        torch = self._torch
        graph = self._graphs.get(key)
        if graph is None:
            launch()  # eager warm-up: loads every kernel before capture
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with graph_capture(torch, graph):
                launch()
            self._graphs[key] = graph
            return
        graph.replay()

    # ------------------------------------------------------------------ public API

    def run_forward(self, store: bool = True) -> None:
        """The forward run from the initial state: drag history and (if ``store``) checkpoints.

        Checkpoint copies and the drag history go through fixed staging buffers outside the
        captured segment, so one graph serves every segment."""
        self.solver._f.copy_(self._start)
        for seg in range(self.n_segments):
            t0 = seg * self.k
            length = min(self.k, self.steps - t0)
            if store:
                self._ckpt[seg].copy_(self.solver._f)
            fields = self._field_hist[t0 : t0 + length] if (store and self.store_fields) else None
            records = self._records[t0 : t0 + length] if fields is not None else None
            self._replay(("fwd", length, store, t0 if fields is not None else 0),
                         lambda: self._forward_core(length, self._drag_stage, fields, records))
            self._drag[t0 : t0 + length].copy_(self._drag_stage[:length])

    def forward(self) -> float:
        """Run forward with checkpoints; returns the loss (host float64)."""
        self.run_forward(store=True)
        drag = self._drag[self.steps - self.window :, 0].double()
        return float(drag.mean().item())

    @property
    def drag_history(self) -> np.ndarray:
        return self._drag[:, 0].double().cpu().numpy()

    def backward(self):
        """dL/dg (d, *shape) and dL/dX (n_markers, d), float64 host arrays.

        Needs the checkpoints of the last ``forward()``."""
        self._lam.zero_()
        self._grad_body.zero_()
        self._grad_w.zero_()
        self._nu.zero_()
        k, n_seg = self.k, self.n_segments
        if n_seg > 1 and (self.steps - (n_seg - 1) * k) % 2:
            raise ValueError("segments after the first must have an even length")

        def length(seg):
            return min(k, self.steps - seg * k)

        def fields(seg, slot):
            if self.store_fields:
                return self._field_hist[seg * k : seg * k + length(seg)]
            return self._field_hist[slot * k : (slot + 1) * k]

        def records(seg):
            return self._records[seg * k : seg * k + length(seg)] if self.store_fields else None

        # The last segment is recomputed on its own; every other segment's recompute rides
        # in the launches of the following segment's backward sweep.
        last = n_seg - 1
        self._segs[0][0].copy_(self._ckpt[last])
        self._replay(("rec", last), lambda: self._recompute(self._segs[0], fields(last, 0), 0, length(last)))
        for i, seg in enumerate(reversed(range(n_seg))):
            slot = i % 2
            partner = None
            if seg > 0:
                self._segs[1 - slot][0].copy_(self._ckpt[seg - 1])
                partner = (self._segs[1 - slot], fields(seg - 1, 1 - slot), length(seg - 1))
            weights = self._body_weight[seg * k : seg * k + length(seg)]
            self._replay(("bwd", seg), lambda: self._sweep(self._segs[slot], fields(seg, slot), records(seg),
                                                           weights, length(seg), partner))
        self._torch.cuda.synchronize()
        p = self.params
        grad_body = self._grad_body.cpu().numpy().reshape(p.d, *p.shape)
        ibm = self.solver._ibm
        grad_x = marker_gradient(ibm.body.markers, ibm.body.stencil_base(), self._grad_w.cpu().numpy())
        return grad_body, grad_x


def drag_loss(problem: DragAdjoint, body_field, markers):
    """Differentiable mean drag: a ``torch.autograd.Function`` over (body_field, markers).

    ``body_field`` holds d * n_nodes values (any shape that reshapes to (d, *shape)) and
    ``markers`` n_markers * d values; both may live on any device and be float32 or float64.
    The loss is a 0-d float64 tensor on ``body_field``'s device. Its backward pass runs
    ``problem.backward()`` on the checkpoints this forward pass stored, so it must run
    before ``problem`` is run forward again. The gradients come back with the shape, dtype
    and device of the inputs they belong to."""
    import torch

    d = problem.params.d

    class DragLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, g, x):
            ctx.meta = ((g.shape, g.dtype, g.device), (x.shape, x.dtype, x.device))
            problem.set_body_force(g.detach())
            problem.set_markers(x.detach().cpu().double().numpy().reshape(-1, d))
            return torch.tensor(problem.forward(), dtype=torch.float64, device=g.device)

        @staticmethod
        def backward(ctx, grad_out):
            gb, gx = problem.backward()
            scale = float(grad_out.detach().cpu())
            out = []
            for grad, (shape, dtype, device) in zip((gb, gx), ctx.meta, strict=True):
                out.append(torch.as_tensor(grad * scale).reshape(shape).to(dtype=dtype, device=device))
            return tuple(out)

    return DragLoss.apply(body_field, markers)
