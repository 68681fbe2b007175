"""drift_amd.solve -- one call that picks exact enumeration, GPU parallel tempering or CPU
parallel tempering by size and hardware (reference: drift/solve.py:24-86).

The result says which method produced it and whether it is certified (only exact enumeration is).
``gpu.gpu_available`` and ``gpu.parallel_tempering_gpu`` are looked up on the module at call time,
so tests can substitute them (reference tests/test_solve.py:79-94).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import gpu
from .solvers.exact import exact_ground_state
from .solvers.parallel_tempering import parallel_tempering


@dataclass
class Solution:
    """A ground-state search result and its provenance."""

    s: np.ndarray
    energy: float
    method: str        # "exact" | "gpu-pt" | "cpu-pt"
    certified: bool    # True only for exact enumeration

    @property
    def n(self) -> int:
        return int(self.s.shape[0])


def solve(
    model,
    *,
    exact_max: int = 18,
    use_gpu: bool = True,
    n_replicas: int | None = None,
    n_rounds: int = 400,
    sweeps_per_round: int = 4,
    seed: int = 0,
    require_certified: bool = False,
) -> Solution:
    """Ground state of ``model``: exact (certified) for n <= exact_max, else GPU parallel
    tempering when the engine is available, else CPU parallel tempering (both uncertified)."""
    n = model.n
    if n <= exact_max:
        s, e, _ = exact_ground_state(model, max_n=max(exact_max, n))
        return Solution(s=s, energy=e, method="exact", certified=True)
    if require_certified:
        raise ValueError(
            f"certified ground state requested, but n={n} > exact_max={exact_max}; the exact "
            "engine enumerates at most ~24 spins.  Shrink the instance, raise exact_max, or pass "
            "require_certified=False to accept an uncertified parallel-tempering minimum.")
    if use_gpu and gpu.gpu_available():
        r = gpu.parallel_tempering_gpu(
            model, n_replicas=n_replicas or 256, n_rounds=n_rounds,
            sweeps_per_round=sweeps_per_round, seed=seed)
        return Solution(s=r.best_s, energy=r.best_E, method="gpu-pt", certified=False)
    r = parallel_tempering(
        model, n_replicas=n_replicas or 32, n_rounds=n_rounds,
        sweeps_per_round=sweeps_per_round, seed=seed)
    return Solution(s=r.best_s, energy=r.best_E, method="cpu-pt", certified=False)
