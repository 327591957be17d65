"""Lattice Boltzmann solver for AMD MI355X (gfx950): hand-written HIP kernels behind the
reference's ``make_solver`` / ``Solver`` API (ShauryaAttreya37/High-Performance-LBM-On-GPU-,
``src/lbm/__init__.py``).

Typical use::

    from lbm import SimulationParams, make_solver

    params = SimulationParams(nx=256, ny=256, tau=0.6, collision="trt")
    solver = make_solver(params, backend="hip-aa")     # or the reference's "cuda-aa"
    solver.initialize(rho=1.0, ux=0.0, uy=0.0)
    solver.run(1000)
    fields = solver.macroscopic()
"""

from lbm.backends import BACKENDS, BackendUnavailable, available_backends, make_solver
from lbm.params import SimulationParams, StabilityWarning
from lbm.solver import SimulationDiverged

__version__ = "0.1.0"

__all__ = [
    "BACKENDS",
    "BackendUnavailable",
    "SimulationDiverged",
    "SimulationParams",
    "StabilityWarning",
    "available_backends",
    "make_solver",
]
