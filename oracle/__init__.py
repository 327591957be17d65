"""CPU oracle for the HIP lattice Boltzmann backend (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg use this package, as the checker or the timed CPU baseline. The product
package ``lbm`` never imports it. See ``reference_solver.py`` for what it
restates and where each part comes from in the reference.
"""

from oracle.reference_solver import ReferenceSolver

__all__ = ["ReferenceSolver"]
