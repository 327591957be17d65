"""HIP backend: fused stream-collide kernels on MI355X, driven through the C ABI.

Replaces ``CudaSolver`` of the reference (``src/lbm/backends/cuda_solver.py:179-325``)
behind the same ``Solver`` contract. Device memory is allocated by PyTorch and
handed to ``liblbm_hip.so`` as raw pointers together with the current torch
stream; no launch synchronises.

Per time step (``pull``/``lds``) or pair of half-steps (``aa``):

1. pre-pass for the layout the next kernel reads (``pull``, ``aa_odd``,
   ``aa_even``): the fused IBM force (``lbm_ibm_force``), then the link rules
   (``lbm_rules_apply_direct``, or ``lbm_rules_apply`` when a rule reads a slot another
   rule writes);
2. the stream-collide kernel (``lbm_pull_step``, ``lbm_aa_odd`` /
   ``lbm_aa_even``, ``lbm_lds_step``).

Between ``run`` calls the AA buffer is in its post-even layout: slot opp(i)
holds f*_i (``state_layout = "aa_odd"``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from lbm.backends import BackendUnavailable, _require_device, graph_capture, load_library
from lbm.backends._native import DTYPE_CODES, FORCE_DENSE, FORCE_FIELD, FORCE_UNIFORM, pointer
from lbm.diagnostics import MacroscopicFields
from lbm.geometry import neighbour_index
from lbm.ibm import rows_are_pulled_moments
from lbm.lattice import VelocitySet
from lbm.link_rules import CompiledAffine, compile_rules, reads_other_targets
from lbm.solver import SimulationDiverged, Solver

VARIANTS = ("pull", "aa", "lds")
TWO_D_ONLY = ("lds",)
KIND_SOLID, KIND_FORCED, KIND_NEAR_SOLID = 1, 2, 4


def check_variant(variant: str, d: int) -> None:
    if variant not in VARIANTS:
        raise ValueError(f"unknown HIP variant {variant!r}; choose from {VARIANTS}")
    if variant in TWO_D_ONLY and d != 2:
        raise ValueError(f"HIP variant {variant!r} is 2D only; use 'pull' or 'aa' for a {d}D grid")


def check_block(block) -> tuple[int, int]:
    bx, by = (int(b) for b in block)
    if bx <= 0 or by <= 0 or bx * by > 1024:
        raise ValueError(f"block must be positive with at most 1024 threads; got {tuple(block)}")
    return bx, by


# D3Q19 launch shape per (variant, dtype): lanes per workgroup and resident workgroups per
# CU, measured at 256^3 (MLUPS against the former common default 256 x 2, cap 2):
#   fp32 AA    256 x 2, cap 2 (16 wavefronts/CU)   37.5 K
#   fp32 pull  256 x 1, cap 4 (16 wavefronts/CU)   38.2 K (was 36.1 K)
#   fp64 both  256 x 1, cap 2 ( 8 wavefronts/CU)   AA 17.6 K (was 16.4 K), pull 17.5 K (17.0 K)
# Fewer concurrent address streams per CU keep the 38 population planes' DRAM pages open
# longer; float64 moves twice the bytes per lane, so half the wavefronts saturate HBM.
_SHAPE_3D = {("aa", "float32"): (512, 2), ("pull", "float32"): (256, 4), ("aa", "float64"): (256, 2),
             ("pull", "float64"): (256, 2)}
# Lattices with solid or forced nodes (the kernels read the kind bytes): every row whose
# ends are solid has a near-solid lane at each end, which takes the longer bounce-back
# path, so smaller workgroups let the rest of the CU move on. 512 x 128 x 128 duct with a
# Bouzidi sphere, AA fp32: 256 x 1 with cap 4 gave 31.9 K MLUPS against 30.7 K for 256 x 2
# with cap 2.
_SHAPE_3D_MASKED = {("aa", "float32"): (256, 4)}
_D2Q9_LANES, _D2Q9_MASKED_LANES = 1024, 512
# Masked D2Q9 lattices up to this many nodes keep rows of at most 512 lanes but stack them
# to 1024 lanes per workgroup: 880 x 164 Bouzidi channel 14.8 -> 15.9 K MLUPS (pull), 512^2
# channel with a cylinder 20.6 -> 25.7 K; from 640^2 to 896^2 the two were within 3 %
# either way, and from 1024^2 up 512 x 1 was as fast or faster.
_D2Q9_SMALL_MASKED_NODES = 2**18


def _masked(params) -> bool:
    return bool(params.solid.any()) or params.immersed is not None


def _shape_3d(params, variant: str) -> tuple[int, int]:
    key = ("aa" if variant == "aa" else "pull", params.dtype)
    return _SHAPE_3D_MASKED.get(key, _SHAPE_3D[key]) if _masked(params) else _SHAPE_3D[key]


def default_block(params, variant: str) -> tuple[int, int]:
    """Whole rows per workgroup where possible. D2Q9: bx covers nx up to 1024 lanes (a
    multiple of the 64-lane wavefront) and by stacks rows up to 1024 lanes; 1024 x 1 was
    1.5-4 % above 256 x 4 on periodic 2048^2 / 4096^2. Lattices with solid or forced nodes
    take rows of up to 512 lanes: Bouzidi/Zou-He channels 4096^2, 8192 x 2048 and 2048^2
    and a 4096^2 lid-driven cavity ran 1-8 % faster at 512 x 1 than at 1024 x 1 (a
    near-solid lane at each end of every row holds its workgroup longer); up to 2^18 nodes
    they stack those rows to 1024 lanes (880 x 164 and 512^2 channels: 7-25 % faster). The LDS-tiled
    kernel keeps bx <= 256: its halo ring costs 2 (bx + 2) + 2 by cells per bx * by, so
    256 x 4 and 128 x 8 win (D2Q9 fp32 4096^2: 76 K MLUPS against 62 K for 64 x 4 and 55 K
    for 32 x 8). D3Q19: bx up to 256 with the measured lanes per workgroup of
    ``_SHAPE_3D`` (``_SHAPE_3D_MASKED`` for lattices with solid or forced nodes).

    A row longer than the widest tile is split into equal tiles (``_even_tiles``): at 1536^2
    D2Q9 two tiles of 768 ran 13-14 % faster than 1024 + 512, at 384^3 D3Q19 two of 192
    10 % faster than 256 + 128, at 320^3 3-5 %, and the 384 x 96 x 96 duct 6 % (pull) and
    13 % (AA)."""
    if params.d == 2:
        masked = _masked(params) and variant != "lds"
        small = params.nx * params.ny <= _D2Q9_SMALL_MASKED_NODES
        lanes = _D2Q9_MASKED_LANES if (masked and not small) else _D2Q9_LANES
        row = 256 if variant == "lds" else (_D2Q9_MASKED_LANES if masked else lanes)
        bx = _even_tiles(params.nx, row)
        return bx, max(1, min(params.ny, lanes // bx))
    bx = _even_tiles(params.nx, 256)
    return bx, max(1, min(params.ny, _shape_3d(params, variant)[0] // bx))


def _even_tiles(nx: int, widest: int) -> int:
    """Tile width for rows of nx nodes: the fewest tiles of at most ``widest`` lanes, each a
    whole number of 64-lane wavefronts and as equal as that allows, so the last tile of a
    row is not left mostly idle (nx = 384, widest 256: 192 + 192 rather than 256 + 128)."""
    tiles = math.ceil(nx / widest)
    return 64 * math.ceil(nx / (64 * tiles))


INFINITY_CACHE_BYTES = 256 * 2**20


def default_groups_per_cu(params, variant: str) -> int:
    """Resident workgroups per compute unit for the stream-collide kernels (0: no cap):
    the measured D3Q19 caps of ``_SHAPE_3D`` / ``_SHAPE_3D_MASKED``. D2Q9 runs uncapped
    except all-fluid lattices whose state is more than twice the 256 MB Infinity Cache,
    where one 1024-lane group per CU moved 4096^2 from 76.2 K to 81.4 K MLUPS (AA, 604 MB) and 81.5 K to 84.8 K (pull,
    1.2 GB); the same cap cost 6-15 % on Bouzidi channels and at 2048^2 (151 / 302 MB)."""
    if params.d == 3:
        return _shape_3d(params, variant)[1]
    if variant == "lds" or _masked(params):
        return 0
    state = params.lattice.q * params.n_nodes * np.dtype(params.dtype).itemsize * (1 if variant == "aa" else 2)
    return 1 if state > 2 * INFINITY_CACHE_BYTES else 0


def node_kinds(lattice: VelocitySet, solid: np.ndarray, forced_nodes=None) -> np.ndarray:
    """Kind byte per node: solid bit, IBM-forced bit, near-solid bit (a solid node at x - c_i)."""
    shape = solid.shape
    flat_solid = solid.ravel()
    kinds = flat_solid.astype(np.uint8) * KIND_SOLID
    if forced_nodes is not None:
        kinds[forced_nodes] |= KIND_FORCED
    if flat_solid.any():
        near = np.zeros(flat_solid.size, dtype=bool)
        nodes = np.arange(flat_solid.size)
        for c in lattice.c:
            if any(c):
                near |= flat_solid[neighbour_index(nodes, [-a for a in c], shape)]
        kinds[near & ~flat_solid] |= KIND_NEAR_SOLID
    return kinds.reshape(shape)


class DeviceArrays:
    """Array namespace of a HIP solver's state: attribute lookups fall through to ``torch``
    (``isfinite``, ``argmax``, ``where``, ...), arrays are created on the solver's device,
    and ``cuda.get_current_stream()`` returns torch's current stream, whose
    ``synchronize()`` the reference's benchmark calls (benchmarks/bench_mlups.py:99-101)."""

    def __init__(self, torch, device):
        self._torch = torch
        self.device = device

        class _Cuda:
            @staticmethod
            def get_current_stream():
                return torch.cuda.current_stream()

            @staticmethod
            def synchronize():
                torch.cuda.synchronize()

        self.cuda = _Cuda()

    def asarray(self, a, dtype=None):
        """A device tensor from a NumPy array, scalar, sequence or tensor (no copy when
        ``a`` already is a device tensor of that dtype)."""
        torch = self._torch
        if dtype is not None and not isinstance(dtype, torch.dtype):
            dtype = getattr(torch, np.dtype(dtype).name)
        if not isinstance(a, torch.Tensor):
            a = np.asarray(a)
        return torch.as_tensor(a, dtype=dtype, device=self.device)

    def __getattr__(self, name):
        return getattr(self._torch, name)

    def __repr__(self) -> str:
        return f"DeviceArrays(torch, device={self.device})"


class _AffinePack:
    """A compiled affine set on the device: term-major slots / coefficients, constants."""

    def __init__(self, torch, compiled: CompiledAffine, dtype, device):
        coeff = np.asarray(compiled.coeff)
        self.n = coeff.shape[0]
        self.m = max(1, coeff.shape[1]) if coeff.ndim == 2 else 1
        slots = np.zeros((self.m, self.n), dtype=np.int64)
        coeffs = np.zeros((self.m, self.n))
        if self.n:
            slots[: coeff.shape[1]] = np.asarray(compiled.slots, dtype=np.int64).T
            coeffs[: coeff.shape[1]] = coeff.T
        self.slot = torch.as_tensor(np.ascontiguousarray(slots), device=device)
        self.coeff = torch.as_tensor(np.ascontiguousarray(coeffs.astype(dtype)), device=device)
        self.constant = torch.as_tensor(np.asarray(compiled.constant, dtype=dtype), device=device)
        self.source = compiled


KIND_RULES = 8
# Nodes a fused launch updates in a workgroup of its own (set only in that launch's copy of
# the kind bytes; the tile workgroups skip them).
KIND_REGION = 16


@dataclass(frozen=True)
class ForcingStep:
    """The one-launch forward step of an IBM body (``lbm_ibm_pull_step``): its kind bytes (the
    stencil nodes carry KIND_REGION) and whether the stencil nodes' lanes sum the moment rows."""

    kind: object
    rows_in_lanes: bool


# The pull step's link rules go inside the step (one launch per step) up to this many nodes;
# above, the saved pre-pass launch is worth less than the rule layer's scattered accesses
# to the end planes of open ducts. Bouzidi body with Zou-He ends, fp32, MLUPS pre-pass ->
# in-step: channels 1024^2 51.9 -> 59.8 K, 2048^2 62.6 -> 65.4 K, 4096^2 75.6 -> 75.8 K;
# ducts 256 x 64 x 64 25.5 -> 29.2 K, 384 x 96 x 96 24.3 -> 24.1 K, 512 x 128 x 128
# 28.9 -> 25.7 K; the 880 x 164 channel 15.5 -> 6.9 us per step.
RULES_IN_STEP_MAX_NODES = 2**22

# Pairs of steps per replayed HIP graph: runs of GRAPH_PAIRS pairs replay one graph, the
# remaining pairs a one-pair graph. Each replay costs the host a graph launch and leaves the
# GPU idle for a few microseconds in between, so longer graphs keep short steps back to back.
GRAPH_PAIRS = 8


def default_rules_in_step(params) -> bool:
    return params.n_nodes <= RULES_IN_STEP_MAX_NODES


def link_rule_index(lattice: VelocitySet, target: np.ndarray, solid: np.ndarray):
    """The rule-node list of a pull-layout rule set for lbm_pull_step_links: the rows sorted
    by (node, direction j), where row r writes slot target[r] = opp(j) n + node (n nodes of
    the grid ``solid``). Returns (order, node_info, node_first): order sorts the rows;
    node_info (count, 4) uint32 = {node, directions i with a solid upstream node x - c_i
    (periodic), directions j with a row, first row} for the nodes with rows, ascending;
    node_first = the first rows and the row count."""
    shape = solid.shape
    n_nodes = int(np.prod(shape))
    target = np.asarray(target, dtype=np.int64)
    if target.size >= 2**31:
        raise ValueError("at most 2^31 - 1 link rules")
    node = target % n_nodes
    direction = np.asarray(lattice.opposite, dtype=np.int64)[target // n_nodes]
    order = np.lexsort((direction, node))
    node, direction = node[order], direction[order]
    if node.size and np.any((node[1:] == node[:-1]) & (direction[1:] == direction[:-1])):
        raise ValueError("two link rules write the same slot")
    starts = np.flatnonzero(np.r_[True, node[1:] != node[:-1]]) if node.size else np.zeros(0, dtype=np.int64)
    bits = np.left_shift(np.uint32(1), direction.astype(np.uint32))
    masks = np.bitwise_or.reduceat(bits, starts) if node.size else np.zeros(0, dtype=np.uint32)
    nodes = node[starts]
    flat_solid = np.asarray(solid).reshape(-1)
    blocked = np.zeros(nodes.size, dtype=np.uint32)
    for i, c in enumerate(lattice.c):
        if any(c):
            upstream = neighbour_index(nodes, [-a for a in c], shape)
            blocked |= flat_solid[upstream].astype(np.uint32) << np.uint32(i)
    if np.any(masks & ~blocked):
        raise ValueError("a link rule sits on a direction whose upstream node is not solid")
    node_info = np.stack([nodes.astype(np.uint32), blocked, masks.astype(np.uint32), starts.astype(np.uint32)], axis=1)
    node_first = np.r_[starts, node.size].astype(np.int32)
    return order, np.ascontiguousarray(node_info), node_first


class HipSolver(Solver):
    """Fused stream-bounce-collide on MI355X.

    Args:
        params: Simulation description.
        variant: ``"pull"``, ``"aa"`` or ``"lds"`` (2D only).
        block: Threads per block (bx, by), at most 1024; default from ``default_block``.
        groups_per_cu: Cap on resident workgroups per compute unit (0: none); default from
            ``default_groups_per_cu``.
        autotune: Time a few block shapes and caps on this grid and keep the fastest.
        graphs: Capture steps (pre-passes and kernels) into HIP graphs and replay them:
            runs of ``2 * GRAPH_PAIRS`` steps from one graph, the remaining pairs from a
            two-step graph.
        rules_in_step: Pull variant with link rules: evaluate them inside the step
            (``lbm_pull_step_links``) rather than in a pre-pass launch (same bits); default
            from ``default_rules_in_step``. The per-step force of ``record_forces`` rides in
            the step whenever the block allows.
    """

    def __init__(self, params, variant: str = "pull", block=None, autotune: bool = False, graphs: bool = False,
                 groups_per_cu=None, rules_in_step=None):
        check_variant(variant, params.d)
        block = default_block(params, variant) if block is None else check_block(block)
        if groups_per_cu is None:
            groups_per_cu = default_groups_per_cu(params, variant)
        if int(groups_per_cu) < 0:
            raise ValueError(f"groups_per_cu must be >= 0; got {groups_per_cu}")
        self._groups_per_cu = int(groups_per_cu)
        super().__init__(params)
        torch = _require_device()
        self._lib = load_library()
        self._torch = torch
        self._xp = None
        self.device = torch.device("cuda")
        self.variant = variant
        self.step_multiple = 2 if variant == "aa" else 1
        self._block = block
        self.graphs = graphs
        self._graphs = {}
        self._forcing_steps = {}
        p = params
        self._np_dtype = np.dtype(p.dtype)
        self._torch_dtype = {"float32": torch.float32, "float64": torch.float64}[p.dtype]
        self._lattice_code = p.lattice.code
        self._dtype_code = DTYPE_CODES[p.dtype]
        self._n = p.n_nodes
        if self._n >= 2**31:
            raise ValueError("grids are limited to 2^31 nodes")
        self._nz = p.nz or 1
        self._rates = p.relaxation_rates
        self._force3 = tuple(p.force) + (0.0,) * (3 - p.d)
        self._body = None

        self._ibm = None
        forced = None
        if p.immersed is not None:
            from lbm.ibm import ImmersedForcing

            self._ibm = ImmersedForcing(p)
            forced = self._ibm.nodes
        kinds = node_kinds(p.lattice, p.solid, forced)
        self._mask_mode = 1 if (p.solid.any() or forced is not None) else 0
        self._kind = torch.as_tensor(kinds.ravel(), device=self.device)
        self._wall = torch.as_tensor(
            np.ascontiguousarray(p.wall_velocity.reshape(p.d, -1).astype(self._np_dtype)), device=self.device
        )
        # walls at rest (the usual case) give exact-zero wall terms: the kernels get no
        # wall array and read no wall velocity
        self._wall_ptr = pointer(self._wall) if np.any(p.wall_velocity) else None
        self._force_mode = FORCE_FIELD if self._ibm is not None else FORCE_UNIFORM
        self._field = None
        if self._ibm is not None:
            self._setup_ibm()

        self._rule_passes = {}
        if p.link_rules is not None and len(p.link_rules):
            layouts = ("aa_odd", "aa_even") if variant == "aa" else ("pull",)
            for layout in layouts:
                compiled = compile_rules(p.lattice, p.link_rules, layout, p.solid, p.wall_velocity)
                pack = _AffinePack(torch, compiled.values, self._np_dtype, self.device)
                pack.target = torch.as_tensor(np.asarray(compiled.target, dtype=np.int64), device=self.device)
                pack.scratch = torch.empty(max(1, pack.n), dtype=self._torch_dtype, device=self.device)
                pack.direct = not reads_other_targets(compiled.target, compiled.values.slots, compiled.values.coeff)
                self._rule_passes[layout] = pack
        # The pull step evaluates its link rules in the consuming lanes (lbm_pull_step_links):
        # the rows sorted by (node, direction) and a per-node index, built once.
        self._links = None
        self.rules_in_step = default_rules_in_step(p) if rules_in_step is None else bool(rules_in_step)
        if variant == "pull" and "pull" in self._rule_passes and self.rules_in_step:
            compiled = compile_rules(p.lattice, p.link_rules, "pull", p.solid, p.wall_velocity)
            order, node_info, node_first = link_rule_index(p.lattice, compiled.target, p.solid)
            v = compiled.values
            links = _AffinePack(torch, CompiledAffine(slots=np.asarray(v.slots)[order],
                                                      coeff=np.asarray(v.coeff)[order],
                                                      constant=np.asarray(v.constant)[order]),
                                self._np_dtype, self.device)
            links.node_info = torch.as_tensor(node_info.view(np.int32), device=self.device)
            links.node_first = torch.as_tensor(node_first, device=self.device)
            links.nodes = len(node_first) - 1
            self._links = links
            flat = kinds.reshape(-1)
            flat[node_info[:, 0].astype(np.int64)] |= KIND_RULES
            self._kind = torch.as_tensor(flat, device=self.device)
        self._affine_cache: dict[int, _AffinePack] = {}
        self._f = torch.empty((p.lattice.q, self._n), dtype=self._torch_dtype, device=self.device)
        self._f_next = None if variant == "aa" else torch.empty_like(self._f)
        self.initialize()
        if autotune:
            from lbm.backends.autotune import tune_block

            self.block, self.groups_per_cu = tune_block(self)

    @property
    def block(self) -> tuple[int, int]:
        return self._block

    @block.setter
    def block(self, value) -> None:
        self._block = check_block(value)
        self._graphs.clear()

    @property
    def groups_per_cu(self) -> int:
        return self._groups_per_cu

    @groups_per_cu.setter
    def groups_per_cu(self, value) -> None:
        if int(value) < 0:
            raise ValueError(f"groups_per_cu must be >= 0; got {value}")
        self._groups_per_cu = int(value)
        self._graphs.clear()

    # ------------------------------------------------------------------ plumbing

    @property
    def _stream(self) -> int:
        return self._torch.cuda.current_stream().cuda_stream

    def _call(self, name: str, *args) -> None:
        self._lib.call(name, *args)

    def _setup_ibm(self) -> None:
        torch, ibm, p = self._torch, self._ibm, self.params
        dev, dt = self.device, self._np_dtype
        self._field = torch.zeros((p.d, self._n), dtype=self._torch_dtype, device=dev)
        self._body_force = torch.zeros(p.d, dtype=self._torch_dtype, device=dev)
        self._ibm_nodes = torch.as_tensor(ibm.nodes, device=dev)
        self._ibm_inverse = torch.as_tensor(ibm.inverse.astype(np.int32), device=dev)
        self._ibm_weights = torch.as_tensor(np.ascontiguousarray(ibm.weights.astype(dt)), device=dev)
        self._ibm_areas = torch.as_tensor(np.asarray(ibm.body.areas, dtype=dt), device=dev)
        self._ibm_start = torch.as_tensor(ibm.spread_start, device=dev)
        self._ibm_entry = torch.as_tensor(ibm.spread_entry, device=dev)
        import ctypes

        size = ctypes.c_int64()
        self._lib.raw("lbm_ibm_scratch_size")(
            self._lattice_code, self._dtype_code, ibm.n_nodes, ibm.n_markers, ctypes.byref(size)
        )
        self._ibm_scratch = torch.empty(size.value // dt.itemsize, dtype=self._torch_dtype, device=dev)
        self._ibm_packs = {}

    def _ibm_pack(self, layout: str) -> _AffinePack:
        if layout not in self._ibm_packs:
            self._ibm_packs[layout] = _AffinePack(self._torch, self._ibm.compiled(layout), self._np_dtype, self.device)
        return self._ibm_packs[layout]

    def set_body_force_field(self, field) -> None:
        """Dense per-node body force (d, *shape) added to the uniform force (``None`` clears it)."""
        if field is None:
            self._body = None
            self._force_mode = FORCE_FIELD if self._ibm is not None else FORCE_UNIFORM
            self._graphs.clear()
            return
        t = self._torch.as_tensor(np.asarray(field), dtype=self._torch_dtype, device=self.device)
        self._body = t.reshape(self.params.d, self._n).contiguous()
        self._force_mode = FORCE_DENSE
        self._graphs.clear()

    # ------------------------------------------------------------------ Solver API

    @property
    def lattice_bytes(self) -> int:
        p = self.params
        return (1 if self.variant == "aa" else 2) * p.lattice.q * self._n * self._np_dtype.itemsize

    @property
    def state_layout(self) -> str:
        return "aa_odd" if self.variant == "aa" else "pull"

    def initialize(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0) -> None:
        drho, u = self._initial_fields(rho, ux, uy, uz)
        torch = self._torch
        dev = self.device
        dr = torch.as_tensor(np.ascontiguousarray(drho).ravel(), device=dev)
        comps = [torch.as_tensor(np.ascontiguousarray(a).ravel(), device=dev) for a in u]
        uz_t = comps[2] if len(comps) == 3 else None
        self._call(
            "lbm_init_equilibrium", self._lattice_code, self._dtype_code, pointer(self._f), pointer(dr),
            pointer(comps[0]), pointer(comps[1]), pointer(uz_t), self._n, int(self.variant == "aa"), self._stream,
        )
        if self._f_next is not None:
            self._f_next.copy_(self._f)
        self._graphs.clear()
        if self._field is not None:
            self._field.zero_()
            self._body_force.zero_()
        self.time_step = 0

    def _step_args(self, stream=None):
        p = self.params
        wp, wm = self._rates
        fx, fy, fz = self._force3
        bx, by = self.block
        return (p.nx, p.ny, self._nz, wp, wm, fx, fy, fz, self._force_mode, self._mask_mode, bx, by,
                self._groups_per_cu, self._stream if stream is None else stream)

    def _advance(self, steps: int) -> None:
        if steps < 0 or steps % self.step_multiple:
            raise ValueError(f"the {self.variant!r} variant advances in multiples of {self.step_multiple} steps; "
                             f"got {steps}")
        if self.graphs and steps >= 2:
            pairs, rest = divmod(steps, 2)
            while pairs:
                k = GRAPH_PAIRS if pairs >= GRAPH_PAIRS else 1
                self._advance_pairs_graphed(k)
                pairs -= k
            if rest:
                self._advance_eager(rest)
            return
        self._advance_eager(steps)

    def _advance_pairs_graphed(self, pairs: int) -> None:
        """2 * ``pairs`` steps from a captured HIP graph (buffers return to their places);
        the first call for a length launches the steps directly and captures them for the
        next. The two-lattice variants keep the graphs per buffer holding the state, since
        an odd number of steps run in between swaps the roles of the two buffers."""
        torch = self._torch
        key = (self._f.data_ptr(), pairs)
        graph = self._graphs.get(key)
        if graph is None:
            self._advance_eager(2 * pairs)
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with graph_capture(torch, graph):
                self._advance_eager(2 * pairs)
            self._graphs[key] = graph
        else:
            graph.replay()

    def _advance_eager(self, steps: int) -> None:
        """Launch ``steps`` steps from a launch list prepared once per call (one per buffer
        parity of the two-lattice variants), so the per-step host work is the launches."""
        stream = self._stream
        sequences = self._step_calls(stream)
        lib = self._lib
        if self.variant == "aa":
            for _ in range(steps // 2):
                lib.call_all(sequences[0])
            return
        for s in range(steps):
            lib.call_all(sequences[s % 2])
        if steps % 2:
            self._f, self._f_next = self._f_next, self._f

    def _step_calls(self, stream) -> list[list[tuple]]:
        """Launch lists of one AA pair, or of one two-lattice step per buffer parity (the
        state in ``_f``, then in ``_f_next``): pre-pass and stream-collide, in order."""
        L, D = self._lattice_code, self._dtype_code
        kind, field, body, wall = pointer(self._kind), pointer(self._field), pointer(self._body), self._wall_ptr
        args = self._step_args(stream)
        if self.variant == "aa":
            f = pointer(self._f)
            return [self._pre_pass_calls("aa_odd", f, stream)
                    + [("lbm_aa_odd", (L, D, f, kind, wall, field, body, *args))]
                    + self._pre_pass_calls("aa_even", f, stream)
                    + [("lbm_aa_even", (L, D, f, kind, field, body, *args))]]
        a, b = pointer(self._f), pointer(self._f_next)
        forcing = self._forcing_in_step()
        if forcing is not None:
            return [[self._ibm_step_call(src, dst, forcing, stream)] for src, dst in ((a, b), (b, a))]
        if self._links is not None and self.block[0] * self.block[1] >= 64:
            return [self._ibm_calls("pull", src, stream) + [self._pull_links_call(src, dst, stream)]
                    for src, dst in ((a, b), (b, a))]
        name = "lbm_lds_step" if self.variant == "lds" else "lbm_pull_step"
        return [self._pre_pass_calls("pull", src, stream) + [(name, (L, D, src, dst, kind, wall, field, body, *args))]
                for src, dst in ((a, b), (b, a))]

    def _forcing_in_step(self, record: bool = False) -> ForcingStep | None:
        """What ``lbm_ibm_pull_step`` needs when the pull variant can run its IBM forcing inside
        the step at the current block (the link rules, if any, in the step; no stencil node
        solid, next to a solid node or a rule node), else None. ``record``: the forcing also
        writes its intermediates."""
        if self.variant != "pull" or self._ibm is None or ("pull" in self._rule_passes and self._links is None):
            return None
        key = (self.block, record, self._kind.data_ptr())
        if key not in self._forcing_steps:
            import ctypes

            ibm, (bx, by) = self._ibm, self.block
            ok = ctypes.c_int()
            self._lib.call("lbm_ibm_pull_step_supported", self._lattice_code, self._dtype_code, ibm.n_nodes,
                           ibm.n_markers, ibm.iterations, int(record), bx, by, ctypes.byref(ok))
            step = None
            blocked = KIND_SOLID | KIND_NEAR_SOLID | KIND_RULES
            if ok.value and not (self._kind.cpu().numpy()[ibm.nodes] & blocked).any():
                kind = self._kind.clone()
                kind[self._ibm_nodes] |= KIND_REGION
                step = ForcingStep(kind, rows_are_pulled_moments(self.params.lattice, ibm.nodes, self.params.shape,
                                                                 ibm.compiled("pull")))
            self._forcing_steps[key] = step
        return self._forcing_steps[key]

    def _ibm_step_call(self, src: int, dst: int, forcing: ForcingStep, stream, field: int | None = None,
                       body_force: int | None = None, record: int | None = None) -> tuple:
        """``lbm_ibm_pull_step`` from ``src`` to ``dst`` (the IBM forcing of ``src``, then the pull
        step with its link rules) as ``_forcing_in_step`` set it up; ``field``, ``body_force``
        (addresses) replace the solver's, ``record`` receives the intermediates."""
        p, ibm, mp = self.params, self._ibm, self._ibm_pack("pull")
        wp, wm = self._rates
        fx, fy, fz = self._force3
        bx, by = self.block
        lk = self._links
        if lk is not None:
            rules = (pointer(lk.node_info), pointer(lk.node_first), lk.nodes, pointer(lk.slot), pointer(lk.coeff),
                     pointer(lk.constant), lk.n, lk.m)
        else:
            rules = (None, None, 0, None, None, None, 0, 1)
        return ("lbm_ibm_pull_step", (
            self._lattice_code, self._dtype_code, src, dst, pointer(forcing.kind), self._wall_ptr,
            pointer(self._field) if field is None else field, pointer(self._body), p.nx, p.ny, self._nz, wp, wm,
            fx, fy, fz, self._force_mode, pointer(mp.slot), pointer(mp.coeff), pointer(mp.constant), mp.m,
            pointer(self._ibm_nodes), ibm.n_nodes, pointer(self._ibm_inverse), pointer(self._ibm_weights),
            pointer(self._ibm_areas), ibm.n_markers, pointer(self._ibm_start), pointer(self._ibm_entry),
            ibm.iterations, pointer(self._body_force) if body_force is None else body_force, record, *rules,
            int(forcing.rows_in_lanes), bx, by, stream,
        ))

    def _pull_links_call(self, src: int, dst: int, stream, force=None, out=None, partial=None,
                         field: int | None = None) -> tuple:
        """The pull step from ``src`` to ``dst`` with the link rules evaluated in the step and,
        when ``force`` (a force pack) is given, the force of the state at ``src`` into
        ``out`` (one reduction group) or ``partial`` (several). ``field`` (an address) replaces
        the solver's IBM force field."""
        p = self.params
        wp, wm = self._rates
        fx, fy, fz = self._force3
        bx, by = self.block
        lk = self._links
        if lk is not None:
            rules = (pointer(lk.node_info), pointer(lk.node_first), lk.nodes, pointer(lk.slot), pointer(lk.coeff),
                     pointer(lk.constant), lk.n, lk.m)
        else:
            rules = (None, None, 0, None, None, None, 0, 1)
        if force is None:
            rows = (None, None, None, None, 0, 1, None, None)
        else:
            rows = (pointer(force.slot), pointer(force.coeff), pointer(force.constant), pointer(force.weight),
                    force.n, force.m, out, partial)
        return ("lbm_pull_step_links", (
            self._lattice_code, self._dtype_code, src, dst, pointer(self._kind), self._wall_ptr,
            pointer(self._field) if field is None else field, pointer(self._body), p.nx, p.ny, self._nz, wp, wm, fx,
            fy, fz, self._force_mode, *rules, *rows, bx, by,
            self._groups_per_cu, stream,
        ))

    def _pre_pass(self, layout: str) -> None:
        """The IBM force field, then the link rules, for the layout the next kernel reads."""
        self._lib.call_all(self._pre_pass_calls(layout, pointer(self._f), self._stream))

    def _pre_pass_calls(self, layout: str, f: int, stream) -> list[tuple]:
        """Launches of the pre-pass on the state at address ``f``: the IBM force field, then
        the link rules. The IBM moments read the state before any rule overwrote a target
        slot, as the NumPy step does (it evaluates the IBM on the stored state,
        src/lbm/solver.py:263-271)."""
        calls = self._ibm_calls(layout, f, stream)
        pack = self._rule_passes.get(layout)
        if pack is not None:
            calls.append(self._rules_call(pack, f, stream))
        return calls

    def _ibm_calls(self, layout: str, f: int, stream) -> list[tuple]:
        """The IBM force-field launch on the state at address ``f`` (none without a body)."""
        calls = []
        if self._ibm is not None:
            ibm = self._ibm
            mp = self._ibm_pack(layout)
            fx, fy, fz = self._force3
            calls.append(("lbm_ibm_force", (
                self._lattice_code, self._dtype_code, f, pointer(mp.slot), pointer(mp.coeff), pointer(mp.constant),
                mp.m, pointer(self._ibm_nodes), ibm.n_nodes, pointer(self._ibm_inverse), pointer(self._ibm_weights),
                pointer(self._ibm_areas), ibm.n_markers, pointer(self._ibm_start), pointer(self._ibm_entry),
                ibm.iterations, fx, fy, fz, self._n, pointer(self._field), pointer(self._body_force),
                pointer(self._ibm_scratch), None, stream,
            )))
        return calls

    def _rules_call(self, pack: _AffinePack, f: int, stream) -> tuple:
        """The launch of one compiled rule set on the state at address ``f``: one launch of
        one rule per lane when no rule reads a slot another rule writes, else the two-phase
        pass (one workgroup up to 4096 rules, else evaluate into scratch, then write)."""
        if pack.direct:
            return ("lbm_rules_apply_direct", (
                self._dtype_code, f, pointer(pack.slot), pointer(pack.coeff), pointer(pack.constant), None,
                pointer(pack.target), pack.n, pack.m, stream,
            ))
        return ("lbm_rules_apply", (
            self._dtype_code, f, pointer(pack.slot), pointer(pack.coeff), pointer(pack.constant),
            pointer(pack.target), pack.n, pack.m, pointer(pack.scratch), stream,
        ))

    def synchronize(self) -> None:
        self._torch.cuda.synchronize()

    def _check_health(self) -> None:
        torch = self._torch
        out = torch.tensor([-1, 0], dtype=torch.int32, device=self.device)
        self._call(
            "lbm_health", self._lattice_code, self._dtype_code, pointer(self._f), pointer(self._kind), self._n,
            int(self.variant == "aa"), pointer(out), self._stream,
        )
        first, nonfinite = out.cpu().numpy().view(np.uint32).tolist()
        if first == 0xFFFFFFFF:
            return
        location = tuple(int(i) for i in np.unravel_index(first, self.params.shape))
        reason = "non-finite populations" if nonfinite else "non-positive density"
        raise SimulationDiverged(self.time_step, location, reason)

    @property
    def xp(self) -> "DeviceArrays":
        """The array namespace of the state (``Solver.xp``, src/lbm/solver.py:74-77): torch
        on the ROCm device, with the CuPy spellings the reference's callers use
        (``xp.cuda.get_current_stream().synchronize()``, ``xp.asarray``)."""
        if self._xp is None:
            self._xp = DeviceArrays(self._torch, self.device)
        return self._xp

    def device_populations(self):
        """Post-collision populations in natural order as a (q, N) device tensor (a copy for AA)."""
        if self.variant == "aa":
            return self._f[list(self.params.lattice.opposite)]
        return self._f

    def _device_populations(self):
        """Post-collision populations in natural direction order on the device, shape
        (q, *shape) (``Solver._device_populations``, src/lbm/solver.py:89-92; for AA a
        permuted copy, as src/lbm/backends/cuda_solver.py:302-305)."""
        p = self.params
        return self.device_populations().reshape(p.lattice.q, *p.shape)

    def populations(self) -> np.ndarray:
        p = self.params
        f = self._f.cpu().numpy()
        if self.variant == "aa":
            f = f[list(p.lattice.opposite)]
        return np.ascontiguousarray(f).reshape(p.lattice.q, *p.shape)

    def _state_buffer(self):
        return self._f

    def evaluate_affine(self, compiled: CompiledAffine) -> np.ndarray:
        pack = self._affine_pack(compiled)
        out = self._torch.empty(max(1, pack.n), dtype=self._torch_dtype, device=self.device)
        if pack.n:
            self._call(
                "lbm_affine_values", self._dtype_code, pointer(self._f), pointer(pack.slot), pointer(pack.coeff),
                pointer(pack.constant), pack.n, pack.m, pointer(out), self._stream,
            )
        return out[: pack.n].cpu().numpy().astype(np.float64)

    def _affine_pack(self, compiled: CompiledAffine) -> _AffinePack:
        key = id(compiled)
        pack = self._affine_cache.get(key)
        if pack is None or pack.source is not compiled:
            pack = _AffinePack(self._torch, compiled, self._np_dtype, self.device)
            self._affine_cache[key] = pack
        return pack

    def _force_pack(self, probe) -> _AffinePack:
        """The probe's rows for the current layout on the device, with their weights and the
        reduction's scratch (uploaded on first use)."""
        pack = self._affine_pack(probe.compiled(self.state_layout))
        if getattr(pack, "weight", None) is None:
            import ctypes

            pack.weight = self._torch.as_tensor(np.ascontiguousarray(probe.weight, dtype=np.float64),
                                                device=self.device)
            size = ctypes.c_int64()
            self._call("lbm_affine_force_scratch_size", pack.n, ctypes.byref(size))
            pack.force_scratch = (self._torch.empty(size.value // 8, dtype=self._torch.float64, device=self.device)
                                  if size.value else None)
        return pack

    def _force_call(self, pack: _AffinePack, f: int, out: int, stream) -> tuple:
        return ("lbm_affine_force", (
            self._dtype_code, f, pointer(pack.slot), pointer(pack.coeff), pointer(pack.constant),
            pointer(pack.weight), pack.n, pack.m, self.params.d, out, pointer(pack.force_scratch), stream,
        ))

    def force_on_device(self, probe, out) -> None:
        """Momentum-exchange force of ``probe`` into the float64 device tensor ``out`` (d,), no host sync."""
        pack = self._force_pack(probe)
        self._lib.call_all([self._force_call(pack, pointer(self._f), pointer(out), self._stream)])

    def record_forces(self, probe, steps: int) -> np.ndarray:
        """Advance ``steps`` steps and return the force after each ``step_multiple`` steps,
        shape (steps // step_multiple, d), float64: the launches of ``_advance`` followed by
        ``force_on_device`` into row k of one device tensor, prepared once; one host copy and
        one health check at the end."""
        m = self.step_multiple
        if steps <= 0 or steps % m:
            raise ValueError(f"steps must be a positive multiple of {m}")
        torch = self._torch
        d = self.params.d
        out = torch.zeros((steps // m, d), dtype=torch.float64, device=self.device)
        pack = self._force_pack(probe)
        stream = self._stream
        lanes = self.block[0] * self.block[1]
        rules_ride = self._links is not None or "pull" not in self._rule_passes
        if self.variant == "pull" and self._mask_mode and rules_ride and lanes % 64 == 0 and lanes <= 1024:
            # (with a rule pre-pass the force must be read before the pre-pass writes the
            # slots it reads, so that case keeps the launches below)
            return self._record_forces_in_step(pack, steps, out)
        sequences = self._step_calls(stream)
        # the state after a step: the AA buffer, or the destination of that parity's step
        states = [pointer(self._f)] if self.variant == "aa" else [pointer(self._f_next), pointer(self._f)]
        base, row = out.data_ptr(), 8 * d
        lib = self._lib
        # Without an immersed body, a step's pre-pass is its link rules alone; when they and the
        # force rows fit one workgroup, the force of step k - 1 and the rules of step k (same
        # state) share one launch (lbm_affine_force_rules), replacing the first call of step k.
        rules = self._rule_passes.get("aa_odd" if self.variant == "aa" else "pull")
        fuse = self._ibm is None and rules is not None and rules.n <= 4096 and pack.n <= 2048
        count = steps // m
        for k in range(count):
            parity = k % len(sequences)
            calls = sequences[parity]
            if fuse and k > 0:
                state = states[(k - 1) % len(sequences)]  # = the source of step k
                calls = [("lbm_affine_force_rules", (
                    self._dtype_code, state, pointer(pack.slot), pointer(pack.coeff), pointer(pack.constant),
                    pointer(pack.weight), pack.n, pack.m, d, base + (k - 1) * row, pointer(rules.slot),
                    pointer(rules.coeff), pointer(rules.constant), pointer(rules.target), rules.n, rules.m, stream,
                ))] + calls[1:]
            if not fuse:
                calls = calls + [self._force_call(pack, states[parity], base + k * row, stream)]
            lib.call_all(calls)
        if fuse:
            lib.call_all([self._force_call(pack, states[(count - 1) % len(sequences)], base + (count - 1) * row,
                                           stream)])
        if self.variant != "aa" and steps % 2:
            self._f, self._f_next = self._f_next, self._f
        self.time_step += steps
        self._check_health()
        return out.cpu().numpy()

    def _record_forces_in_step(self, pack: _AffinePack, steps: int, out) -> np.ndarray:
        """``record_forces`` for the pull variant: the force of the state a step reads is
        reduced by the extra workgroup layer of that step's own launch
        (lbm_pull_step_links), so a step with its force is one launch (plus the IBM
        pre-pass when there is a body). Step k records the force after step k - 1; the
        force after the last step is one lbm_affine_force launch. Above 2048 rows the
        layer writes each step's partial sums, added up per row by one lbm_force_totals
        launch at the end. Every row is summed in lbm_affine_force's order, bit for bit."""
        import ctypes

        torch, lib, stream = self._torch, self._lib, self._stream
        d = self.params.d
        groups = ctypes.c_int()
        self._call("lbm_affine_force_groups", pack.n, ctypes.byref(groups))
        groups = groups.value
        partial = (torch.empty((steps - 1, groups, 3), dtype=torch.float64, device=self.device)
                   if groups > 1 and steps > 1 else None)
        a, b = pointer(self._f), pointer(self._f_next)
        base, row = out.data_ptr(), 8 * d
        prow = 8 * 3 * groups
        for k in range(steps):
            src, dst = (a, b) if k % 2 == 0 else (b, a)
            if k == 0:
                step = self._pull_links_call(src, dst, stream)
            elif partial is None:
                step = self._pull_links_call(src, dst, stream, pack, base + (k - 1) * row, None)
            else:
                step = self._pull_links_call(src, dst, stream, pack, None, partial.data_ptr() + (k - 1) * prow)
            lib.call_all(self._ibm_calls("pull", src, stream) + [step])
        final = b if steps % 2 else a
        lib.call_all([self._force_call(pack, final, base + (steps - 1) * row, stream)])
        if partial is not None:
            self._call("lbm_force_totals", partial.data_ptr(), steps - 1, groups, d, base, stream)
        if steps % 2:
            self._f, self._f_next = self._f_next, self._f
        self.time_step += steps
        self._check_health()
        return out.cpu().numpy()

    @property
    def immersed_force(self) -> np.ndarray:
        if self._ibm is None:
            raise AttributeError("this simulation has no immersed body")
        return self._body_force.cpu().numpy().astype(np.float64)

    def macroscopic(self) -> MacroscopicFields:
        p = self.params
        torch = self._torch
        rho = torch.empty(self._n, dtype=self._torch_dtype, device=self.device)
        u = torch.empty((p.d, self._n), dtype=self._torch_dtype, device=self.device)
        fx, fy, fz = self._force3
        self._call(
            "lbm_macroscopic", self._lattice_code, self._dtype_code, pointer(self._f), pointer(self._kind),
            pointer(self._field), self._n, int(self.variant == "aa"), fx, fy, fz,
            FORCE_FIELD if self._field is not None else FORCE_UNIFORM, pointer(rho), pointer(u), self._stream,
        )
        rho_h = rho.cpu().numpy().reshape(p.shape)
        u_h = u.cpu().numpy().reshape(p.d, *p.shape)
        return MacroscopicFields(rho=rho_h, u=tuple(u_h[a].copy() for a in range(p.d)), solid=np.array(p.solid))

    def kernel_resources(self) -> dict[str, dict[str, int]]:
        """Registers, scratch bytes per lane, static LDS bytes and max threads of each step kernel."""
        import ctypes

        values = (ctypes.c_int * (4 * 16))()
        names = ctypes.create_string_buffer(1024)
        count = ctypes.c_int()
        self._call("lbm_kernel_resources", self._lattice_code, self._dtype_code, values, 16, names, 1024,
                   ctypes.byref(count))
        out = {}
        for k, name in enumerate(names.value.decode().split(";")[: count.value]):
            out[name] = {
                "registers": values[4 * k],
                "local_bytes": values[4 * k + 1],
                "shared_bytes": values[4 * k + 2],
                "max_threads": values[4 * k + 3],
            }
        return out


__all__ = ["HipSolver", "DeviceArrays", "KIND_RULES", "KIND_REGION", "link_rule_index", "default_rules_in_step", "VARIANTS", "TWO_D_ONLY", "BackendUnavailable", "node_kinds", "default_block",
           "default_groups_per_cu"]
