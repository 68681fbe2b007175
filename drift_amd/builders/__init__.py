"""Instance builders for the parallel-tempering path: the reference's dense generators (whose
seeded output is pinned by the reference bank's fingerprints) and sparse generators for the large
lattice, random-graph and batched workloads."""

from .crystal import column_period, crystal_2d, is_striped
from .lattice import (
    cubic_pmj_glass,
    er_sparse_weighted,
    random_regular_gaussian,
    square_ferromagnet,
)
from .qubo import cut_value, maxcut_ising, qubo_to_ising, random_graph

__all__ = [
    "column_period",
    "crystal_2d",
    "cubic_pmj_glass",
    "cut_value",
    "er_sparse_weighted",
    "is_striped",
    "maxcut_ising",
    "qubo_to_ising",
    "random_graph",
    "random_regular_gaussian",
    "square_ferromagnet",
]
