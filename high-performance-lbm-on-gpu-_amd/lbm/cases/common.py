"""Helpers shared by the validation cases (the roles of the reference's
``src/lbm/cases/common.py:16-60``: steady-state loop, even step counts, backend options)."""

from __future__ import annotations

import math
from collections.abc import Callable, Iterator

import numpy as np

from lbm.backends import ALIASES, BACKENDS


def _changes(solver, observe: Callable, interval: int, max_steps: int) -> Iterator[float]:
    """Largest absolute change of ``observe`` over each further ``interval`` steps."""
    before = observe(solver)
    while solver.time_step < max_steps:
        solver.run(interval, check_every=interval)
        after = observe(solver)
        yield float(np.abs(after - before).max())
        before = after


def run_until_steady(solver, observe: Callable, *, interval: int, tol: float, max_steps: int) -> int:
    """Advance in chunks of ``interval`` steps until the observed array changes by less than
    ``tol`` (largest absolute difference between consecutive checks); returns the step count.
    ``RuntimeError`` when ``max_steps`` pass first."""
    change = math.inf
    for change in _changes(solver, observe, interval, max_steps):
        if change < tol:
            return solver.time_step
    raise RuntimeError(f"no steady state after {max_steps} steps: last change {change:.3e} >= tol {tol:.1e}")


def even(steps: float) -> int:
    """Nearest even integer, at least 2 (the AA pattern advances in pairs)."""
    return 2 * max(1, round(steps / 2))


_BACKEND_OPTIONS = (
    ("--backend", "hip-aa", BACKENDS + tuple(ALIASES)),
    ("--dtype", "float64", ("float32", "float64")),
    ("--collision", "trt", ("bgk", "trt")),
)


def add_backend_arguments(parser) -> None:
    """--backend, --dtype and --collision options of the case command lines."""
    for flag, default, choices in _BACKEND_OPTIONS:
        parser.add_argument(flag, default=default, choices=choices)
