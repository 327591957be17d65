"""Helpers shared by the validation cases (``src/lbm/cases/common.py:16-60`` of the reference)."""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from lbm.backends import ALIASES, BACKENDS


def run_until_steady(solver, observe: Callable, *, interval: int, tol: float, max_steps: int) -> int:
    """Advance in chunks of ``interval`` until ``observe`` changes by less than ``tol`` (max abs)."""
    previous = observe(solver)
    change = np.inf
    while solver.time_step < max_steps:
        solver.run(interval, check_every=interval)
        current = observe(solver)
        change = float(np.max(np.abs(current - previous)))
        if change < tol:
            return solver.time_step
        previous = current
    raise RuntimeError(f"no steady state after {max_steps} steps: last change {change:.3e} >= tol {tol:.1e}")


def even(steps: float) -> int:
    """Nearest even integer >= 2 (AA advances in pairs)."""
    return max(2, 2 * round(steps / 2))


def add_backend_arguments(parser) -> None:
    parser.add_argument("--backend", default="hip-aa", choices=BACKENDS + tuple(ALIASES))
    parser.add_argument("--dtype", default="float64", choices=("float32", "float64"))
    parser.add_argument("--collision", default="trt", choices=("bgk", "trt"))
