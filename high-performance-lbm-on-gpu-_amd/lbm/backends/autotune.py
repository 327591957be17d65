"""Project-written autotuner: pick workgroup shapes (and persistent grid sizes) by timing.

Candidates are timed on the solver's own grid with HIP events on the launch stream,
after warm-up, with the simulation state saved before and restored after, so tuning
never changes results (every shape computes the same bits). Choices are cached per
(kind, lattice, dtype, grid sizes) for the life of the process.
"""

from __future__ import annotations

BLOCKS = ((256, 1), (256, 2), (256, 4), (128, 2), (128, 4), (128, 8), (64, 4), (64, 8), (64, 16), (512, 1),
          (512, 2), (1024, 1))
GROUPS_PER_CU = (2, 4, 8, 16, 32)
BATCHED_BLOCKS = ((256, 1), (512, 1), (1024, 1))
"""Work-item sizes of the batched step (a work item is a run of consecutive nodes, so only
the thread count of the workgroup matters)."""
_CACHE: dict = {}


def _time(torch, fn, reps: int) -> float:
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    start.record(stream)
    for _ in range(reps):
        fn()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def candidate_blocks(nx: int, variant: str = "pull"):
    """Shapes whose x extent does not exceed the row much (no mostly idle wavefronts)."""
    if variant == "lds":
        return [(256, 4), (128, 8), (64, 16), (512, 2), (128, 4), (64, 8)]
    limit = max(64, 64 * -(-nx // 64))
    return [b for b in BLOCKS if b[0] <= limit]


SOLVER_CAPS = (0, 1, 2, 3, 4)
"""Resident-workgroup caps per compute unit tried with every block shape (0: none)."""


def tune_block(solver, candidates=None, caps=SOLVER_CAPS, reps: int = 6) -> tuple[tuple[int, int], int]:
    """Fastest (block, groups_per_cu) for a ``HipSolver``; its state is restored afterwards."""
    p = solver.params
    key = ("solver", solver.variant, p.lattice.name, p.dtype, p.nx, p.ny, p.nz, solver._mask_mode)
    if key in _CACHE:
        return _CACHE[key]
    torch = solver._torch
    saved = solver._f.clone()
    saved_next = None if solver._f_next is None else solver._f_next.clone()
    steps = solver.step_multiple
    best, best_ms = (solver.block, solver.groups_per_cu), float("inf")
    for block in candidates or candidate_blocks(p.nx, solver.variant):
        for cap in caps:
            solver.block, solver.groups_per_cu = block, cap
            try:
                ms = _time(torch, lambda: solver._advance(steps), reps)
            except (ValueError, RuntimeError):
                continue
            if ms < best_ms:
                best, best_ms = (block, cap), ms
    solver._f.copy_(saved)
    if saved_next is not None:
        solver._f_next.copy_(saved_next)
    solver.block, solver.groups_per_cu = best
    _CACHE[key] = best
    return best


def tune_batched(batched, reps: int = 4, caps=(0, 1, 2, 4)):
    """Fastest (block, groups, groups_per_cu) for a ``BatchedSolver``; its state is restored
    afterwards."""
    key = ("batched", batched.lattice.name, batched.dtype, tuple(batched.nodes))
    if key in _CACHE:
        return _CACHE[key]
    torch = batched._torch
    saved = batched.snapshot()
    graphs = batched.graphs
    batched.graphs = False
    best, best_ms = (batched.block, batched.groups, batched.groups_per_cu), float("inf")
    for block in BATCHED_BLOCKS:
        batched.block = block
        batched._build_tiles()
        for per_cu in GROUPS_PER_CU:
            batched.groups = per_cu * 256
            for cap in caps:
                batched.groups_per_cu = cap
                try:
                    ms = _time(torch, batched._pair, reps)
                except (ValueError, RuntimeError):  # a shape the kernels were not compiled for
                    continue
                if ms < best_ms:
                    best, best_ms = (block, batched.groups, cap), ms
    batched.graphs = graphs
    batched.block, batched.groups, batched.groups_per_cu = best
    batched._build_tiles()
    batched.restore(saved)
    _CACHE[key] = best
    return best
