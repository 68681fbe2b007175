"""Solvers on the parallel-tempering path: the CPU parallel-tempering routine the dispatcher falls
back to without a GPU, and exact enumeration for small n."""

from .exact import all_configs, exact_ground_state, exact_thermal
from .parallel_tempering import PtResult, geometric_ladder, parallel_tempering

__all__ = [
    "PtResult",
    "all_configs",
    "exact_ground_state",
    "exact_thermal",
    "geometric_ladder",
    "parallel_tempering",
]
