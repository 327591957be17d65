"""Backend registry and factory (``src/lbm/backends/__init__.py:23-83`` of the reference).

==============  ==================================================================
``hip-pull``    Fused pull stream-bounce-collide, two lattices (ping-pong).
``hip-aa``      AA pattern: one lattice updated in place, half the memory.
``hip-lds``     Two lattices with each block's tile + halo staged in LDS (2D only).
==============  ==================================================================

The reference's GPU names map onto these: ``cuda-naive`` -> ``hip-pull``,
``cuda-aa`` -> ``hip-aa``, ``cuda-tiled`` -> ``hip-lds``, so callers written
against the reference keep working. The reference's ``numpy`` / ``cupy``
backends are CPU / unfused array code; this package ships only the HIP path.
The NumPy restatement of the reference solver lives in the repository's
``oracle/`` directory as the correctness oracle of the test suite.
"""

from __future__ import annotations

import contextlib
import functools
import gc

BACKENDS = ("hip-pull", "hip-aa", "hip-lds")
ALIASES = {"cuda-naive": "hip-pull", "cuda-aa": "hip-aa", "cuda-tiled": "hip-lds"}


class BackendUnavailable(RuntimeError):
    """The HIP library or a ROCm device is missing."""


def _require_device():
    try:
        import torch
    except ImportError as exc:  # pragma: no cover - torch is part of the image
        raise BackendUnavailable("PyTorch is not installed") from exc
    if not torch.cuda.is_available():
        raise BackendUnavailable("no ROCm GPU is visible (torch.cuda.is_available() is False)")
    return torch


@contextlib.contextmanager
def graph_capture(torch, graph):
    """``torch.cuda.graph(graph)`` with Python's cyclic garbage collector off until the
    capture has ended. ``torch.cuda.graph`` collects garbage before it starts capturing; a
    collection that an allocation triggers inside the capture can run the destructor of
    another object's HIP graph, a call that stream capture does not allow, and the process
    aborts."""
    enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.graph(graph):
            yield
    finally:
        if enabled:
            gc.enable()


def load_library():
    """The loaded ``liblbm_hip.so``; ``BackendUnavailable`` if it is missing."""
    from lbm.backends._native import library

    try:
        return library()
    except (OSError, FileNotFoundError) as exc:
        raise BackendUnavailable(f"liblbm_hip.so cannot be loaded: {exc}") from exc


@functools.cache
def gpu_available() -> bool:
    try:
        _require_device()
        load_library()
    except BackendUnavailable:
        return False
    return True


def available_backends() -> tuple[str, ...]:
    return BACKENDS if gpu_available() else ()


def make_solver(params, backend: str = "hip-aa", **options):
    """Create a HIP solver for ``params``.

    Args:
        params: ``SimulationParams``.
        backend: One of ``BACKENDS`` or a reference alias (``cuda-naive``, ``cuda-aa``, ``cuda-tiled``).
        **options: Passed to ``HipSolver`` (``block=(bx, by)``, ``autotune=True``).

    Raises:
        ValueError: Unknown backend name.
        BackendUnavailable: No ROCm device or the library is not built.
    """
    name = ALIASES.get(backend, backend)
    if name in ("numpy", "cupy"):
        raise BackendUnavailable(
            f"backend {backend!r} is CPU/array code of the reference; this package runs on HIP only "
            f"(choose from {BACKENDS})"
        )
    if name not in BACKENDS:
        raise ValueError(f"unknown backend {backend!r}; choose from {BACKENDS + tuple(ALIASES)}")
    from lbm.backends.hip_solver import HipSolver

    return HipSolver(params, variant=name.removeprefix("hip-"), **options)
