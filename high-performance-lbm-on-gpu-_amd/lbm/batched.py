"""Many independent lattices of uneven sizes advanced together on one GPU.

No counterpart in the reference (one lattice per solver). Every lattice is a pull-layout
simulation (``SimulationParams``: walls, link rules, uniform force; no immersed body)
of the same velocity set and precision. All lattices share three device pools
(populations x2, kind bytes, wall velocities); a step is

1. one rule pre-pass over the concatenated rules of every lattice, rows keeping their
   own padded width so they round as alone (``lbm_rules_apply_direct``, one launch, when
   no rule reads a slot another rule writes; else ``lbm_rules_apply_widths``);
2. one persistent launch of ``lbm_batched_pull_step`` over the (lattice, tile) work table.

Runs of steps (an even number, so the pools return to their places) are captured in HIP
graphs and replayed: 2 * GRAPH_PAIRS steps per graph, the remaining pairs from a two-step
graph. Each lattice's
populations equal, bit for bit, what ``HipSolver(params, "pull")`` computes for it alone
(``tests/test_batched.py``). Block shape and the number of persistent workgroups come from
``lbm.backends.autotune.tune_batched`` when ``autotune=True``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from lbm.backends import _require_device, graph_capture, load_library
from lbm.backends._native import DTYPE_CODES, pointer
from lbm.backends.hip_solver import GRAPH_PAIRS, check_block, node_kinds
from lbm.link_rules import compile_rules, reads_other_targets

DEFAULT_GROUPS_PER_CU = 8
COMPUTE_UNITS = 256
PLANE_ALIGN = 64
"""Population planes of each lattice start every PLANE_ALIGN elements in the pool (the
node count rounded up, as ``lbm_batched_describe`` lays them out): every plane begins on a
cache-line boundary."""


def plane_stride(n: int) -> int:
    return -(-int(n) // PLANE_ALIGN) * PLANE_ALIGN


class BatchedSolver:
    """Pull-layout lattices ``params_list`` stepped by one persistent launch per step.

    Args:
        params_list: ``SimulationParams`` of every lattice (same lattice and dtype, no IBM).
        block: workgroup shape (bx, by); a work item is bx * by consecutive nodes of one
            lattice; default (1024, 1), (256, 2) for float64 D3Q19 (512-lane kernels).
        groups: persistent workgroups; default 8 per compute unit.
        graphs: capture the two-step cycle into a HIP graph.
        autotune: pick block and groups by timing (``lbm.backends.autotune``).
    """

    def __init__(self, params_list, block=None, groups=None, graphs: bool = True, autotune: bool = False):
        params_list = list(params_list)
        if not params_list:
            raise ValueError("need at least one lattice")
        first = params_list[0]
        for p in params_list:
            if p.lattice.name != first.lattice.name or p.dtype != first.dtype:
                raise ValueError("batched lattices must share the velocity set and the dtype")
            if p.immersed is not None:
                raise ValueError("batched lattices do not support immersed bodies")
        torch = _require_device()
        self._torch = torch
        self._lib = load_library()
        self.params_list = params_list
        self.lattice = first.lattice
        self.dtype = first.dtype
        self.device = torch.device("cuda")
        self._np_dtype = np.dtype(first.dtype)
        self._tdt = {"float32": torch.float32, "float64": torch.float64}[first.dtype]
        self._lcode, self._dcode = first.lattice.code, DTYPE_CODES[first.dtype]
        q, d = first.lattice.q, first.d
        self.nodes = [p.n_nodes for p in params_list]
        self.total_nodes = int(sum(self.nodes))
        if block is None:  # float64 D3Q19 kernels are compiled for at most 512 lanes
            block = (256, 2) if (d == 3 and first.dtype == "float64") else (1024, 1)
        self.block = check_block(block)
        self.groups = int(groups) if groups else DEFAULT_GROUPS_PER_CU * COMPUTE_UNITS
        self.groups_per_cu = 0  # cap on resident workgroups per CU (0: none); set by the autotuner
        self.graphs = graphs
        self.time_step = 0

        n_l = len(params_list)
        size = ctypes.c_int64()
        self._lib.call("lbm_batched_desc_size", self._dcode, ctypes.byref(size))
        host = np.zeros(n_l * size.value, dtype=np.uint8)
        ints = lambda vals: np.ascontiguousarray(vals, dtype=np.int32)  # noqa: E731
        nx = ints([p.nx for p in params_list])
        ny = ints([p.ny for p in params_list])
        nz = ints([p.nz or 1 for p in params_list])
        rates = np.array([p.relaxation_rates for p in params_list], dtype=np.float64)
        wp, wm = np.ascontiguousarray(rates[:, 0]), np.ascontiguousarray(rates[:, 1])
        force = np.ascontiguousarray(np.array([tuple(p.force) + (0.0,) * (3 - d) for p in params_list]))
        self._lib.call("lbm_batched_describe", self._lcode, self._dcode, n_l, nx.ctypes.data, ny.ctypes.data,
                       nz.ctypes.data, wp.ctypes.data, wm.ctypes.data, force.ctypes.data, host.ctypes.data)
        self._desc = torch.as_tensor(host, device=self.device)
        self.strides = [plane_stride(n) for n in self.nodes]
        self.f_offset = np.concatenate([[0], np.cumsum([q * s for s in self.strides])]).astype(np.int64)
        node_offset = np.concatenate([[0], np.cumsum(self.nodes)]).astype(np.int64)
        self.node_offset = node_offset

        kinds = np.concatenate([node_kinds(p.lattice, p.solid).ravel() for p in params_list])
        walls = np.concatenate([p.wall_velocity.reshape(-1).astype(self._np_dtype) for p in params_list])
        self._kind = torch.as_tensor(kinds, device=self.device)
        self._wall = torch.as_tensor(walls, device=self.device)
        self._wall_ptr = pointer(self._wall) if np.any(walls) else None  # none moving: no wall reads
        self._mask_mode = 1 if any(p.solid.any() for p in params_list) else 0
        self._fa = torch.zeros(int(self.f_offset[-1]), dtype=self._tdt, device=self.device)
        self._fb = torch.zeros_like(self._fa)
        self._build_rules()
        self._build_tiles()
        self._graphs = {}
        for i in range(n_l):
            self.initialize(i)
        if autotune:
            from lbm.backends.autotune import tune_batched

            self.block, self.groups, self.groups_per_cu = tune_batched(self)
            self._build_tiles()

    # ------------------------------------------------------------------ set-up

    def _build_rules(self) -> None:
        slots, coeffs, consts, targets, widths = [], [], [], [], []
        for i, p in enumerate(self.params_list):
            if p.link_rules is None or not len(p.link_rules):
                continue
            c = compile_rules(p.lattice, p.link_rules, "pull", p.solid, p.wall_velocity)
            base, n, stride = int(self.f_offset[i]), p.n_nodes, self.strides[i]

            def pooled(flat, base=base, n=n, stride=stride):  # slot (k, y) = k n + y -> k stride + y
                flat = np.asarray(flat, dtype=np.int64)
                return flat // n * stride + flat % n + base

            slots.append(pooled(c.values.slots))
            coeffs.append(np.asarray(c.values.coeff))
            consts.append(np.asarray(c.values.constant))
            targets.append(pooled(c.target))
            widths.append(np.full(len(c.target), c.values.coeff.shape[1], dtype=np.int32))
        self._n_rules = int(sum(len(t) for t in targets))
        if not self._n_rules:
            return
        m = max(s.shape[1] for s in slots)
        pad = lambda a: np.pad(a, ((0, 0), (0, m - a.shape[1])))  # noqa: E731
        slot = np.concatenate([pad(s) for s in slots]).T
        coeff = np.concatenate([pad(c) for c in coeffs]).T.astype(self._np_dtype)
        torch, dev = self._torch, self.device
        self._rule_m = m
        self._rule_slot = torch.as_tensor(np.ascontiguousarray(slot), device=dev)
        self._rule_coeff = torch.as_tensor(np.ascontiguousarray(coeff), device=dev)
        self._rule_const = torch.as_tensor(np.concatenate(consts).astype(self._np_dtype), device=dev)
        self._rule_target = torch.as_tensor(np.concatenate(targets), device=dev)
        self._rule_width = torch.as_tensor(np.concatenate(widths), device=dev)
        self._rule_scratch = torch.empty(self._n_rules, dtype=self._tdt, device=dev)
        # lattices share no slots, so the concatenation is one-launch safe when each set is
        self._rule_direct = not reads_other_targets(np.concatenate(targets), slot.T, coeff.T)

    def _build_tiles(self) -> None:
        """Work items (lattice, first node): runs of bx * by consecutive nodes of one lattice."""
        bx, by = self.block
        size = bx * by
        parts = []
        for i, n in enumerate(self.nodes):
            starts = np.arange(0, n, size, dtype=np.int64)
            parts.append(np.stack([np.full(starts.size, i, dtype=np.int64), starts], axis=1))
        tiles = np.concatenate(parts).astype(np.int32)
        self._tiles = self._torch.as_tensor(tiles.ravel(), device=self.device)
        self._n_tiles = len(tiles)
        self._graphs = {}

    # ------------------------------------------------------------------ state

    def view(self, i: int, pool=None):
        """Populations of lattice i as a (q, N_i) view of the current pool (planes
        ``strides[i]`` elements apart)."""
        pool = self._fa if pool is None else pool
        q = self.lattice.q
        return pool[int(self.f_offset[i]) : int(self.f_offset[i + 1])].view(q, self.strides[i])[:, : self.nodes[i]]

    def initialize(self, i: int, rho=1.0, ux=0.0, uy=0.0, uz=0.0) -> None:
        """Equilibrium of lattice i (same construction as ``Solver.initialize``)."""
        from lbm.solver import initial_fields

        p = self.params_list[i]
        drho, u = initial_fields(p, rho, ux, uy, uz)
        torch, dev = self._torch, self.device
        dr = torch.as_tensor(np.ascontiguousarray(drho).ravel(), device=dev)
        comps = [torch.as_tensor(np.ascontiguousarray(a).ravel(), device=dev) for a in u]
        f = torch.empty((p.lattice.q, p.n_nodes), dtype=self._tdt, device=dev)  # contiguous planes
        self._lib.call("lbm_init_equilibrium", self._lcode, self._dcode, pointer(f), pointer(dr), pointer(comps[0]),
                       pointer(comps[1]), pointer(comps[2] if len(comps) == 3 else None), p.n_nodes, 0,
                       torch.cuda.current_stream().cuda_stream)
        self.view(i).copy_(f)
        self.view(i, self._fb).copy_(f)

    def _step(self, src, dst) -> None:
        stream = self._torch.cuda.current_stream().cuda_stream
        if self._n_rules and self._rule_direct:
            self._lib.call("lbm_rules_apply_direct", self._dcode, pointer(src), pointer(self._rule_slot),
                           pointer(self._rule_coeff), pointer(self._rule_const), pointer(self._rule_width),
                           pointer(self._rule_target), self._n_rules, self._rule_m, stream)
        elif self._n_rules:
            self._lib.call("lbm_rules_apply_widths", self._dcode, pointer(src), pointer(self._rule_slot),
                           pointer(self._rule_coeff), pointer(self._rule_const), pointer(self._rule_target),
                           pointer(self._rule_width), self._n_rules, self._rule_m, pointer(self._rule_scratch), stream)
        bx, by = self.block
        self._lib.call("lbm_batched_pull_step", self._lcode, self._dcode, pointer(src), pointer(dst),
                       pointer(self._kind), self._wall_ptr, pointer(self._desc), pointer(self._tiles),
                       self._n_tiles, self._mask_mode, bx, by, self.groups, self.groups_per_cu, stream)

    def _pair(self) -> None:
        self._step(self._fa, self._fb)
        self._step(self._fb, self._fa)

    def run(self, steps: int) -> None:
        """Advance every lattice by ``steps`` steps: runs of GRAPH_PAIRS pairs replayed from one
        captured graph, the remaining pairs from a one-pair graph (the first run of each
        length is launched directly and captured for the next), an odd last step launched."""
        torch = self._torch
        pairs, odd = divmod(int(steps), 2)
        while pairs:
            k = GRAPH_PAIRS if pairs >= GRAPH_PAIRS else 1
            pairs -= k
            if not self.graphs:
                for _ in range(k):
                    self._pair()
                continue
            # the launches carry the grid size and the occupancy cap
            key = (k, self.groups, self.groups_per_cu)
            graph = self._graphs.get(key)
            if graph is None:
                for _ in range(k):
                    self._pair()
                torch.cuda.synchronize()
                graph = torch.cuda.CUDAGraph()
                with graph_capture(torch, graph):
                    for _ in range(k):
                        self._pair()
                self._graphs[key] = graph
            else:
                graph.replay()
        if odd:
            self._step(self._fa, self._fb)
            self._fa, self._fb = self._fb, self._fa
            self._graphs = {}
        self.time_step += int(steps)

    def synchronize(self) -> None:
        self._torch.cuda.synchronize()

    def populations(self, i: int) -> np.ndarray:
        p = self.params_list[i]
        return self.view(i).cpu().numpy().reshape(p.lattice.q, *p.shape)

    def snapshot(self):
        return self._fa.clone()

    def restore(self, state) -> None:
        self._fa.copy_(state)
        self._fb.copy_(state)
