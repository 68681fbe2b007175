"""drift_amd -- a gfx950 (MI355X) parallel-tempering engine behind DRiFT's Ising/QUBO interface.

Mirrors the reference's hot-path surface: ``IsingModel`` (drift/ising.py), ``PtResult`` /
``geometric_ladder`` / ``parallel_tempering`` (drift/solvers/parallel_tempering.py),
``exact_ground_state`` (drift/solvers/exact.py), ``solve`` (drift/solve.py) and the GPU entry point
``drift_amd.gpu.parallel_tempering_gpu`` (drift/gpu.py).  As in the reference, ``gpu`` is not
re-exported here; import it explicitly.
"""

from .ising import IsingModel, SparseIsingModel
from .solve import Solution, solve
from .solvers.exact import exact_ground_state, exact_thermal
from .solvers.parallel_tempering import PtResult, geometric_ladder, parallel_tempering

__all__ = [
    "IsingModel",
    "PtResult",
    "Solution",
    "SparseIsingModel",
    "exact_ground_state",
    "exact_thermal",
    "geometric_ladder",
    "parallel_tempering",
    "solve",
]
