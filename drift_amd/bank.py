"""Instance-bank rows and the large-n quality checks on the GPU path.

The reference measures its solver on a bank of fingerprinted instances
(drift/benchmarks/bank.py:120-206) through a runner that calls ``solve`` once per row
(experiments/scale_sweep.py:141-213), and checks solution quality past the exact wall on two
instances with known optima: planted bipartite max-cut, whose optimum cuts every edge
(experiments/phase14_gpu.py:88-99), and the frustrated 2-D crystal, whose ground state is the
period-4 stripe with E/n = -2 (experiments/crystal_gpu.py:24-31).  This module is that runner and
those checks for this package's ``solve`` (exact below ``exact_max`` spins, then the gfx950 engine).
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .builders.crystal import column_period, crystal_2d, is_striped
from .builders.families import build_model, dense_fingerprint
from .builders.qubo import cut_value, maxcut_ising
from .solve import solve

__all__ = ["BankRow", "bank_model", "sweep_row", "planted_bipartite", "bipartite_check",
           "crystal_check"]


@dataclass
class BankRow:
    """One solved bank instance: the reference's SweepRow fields (experiments/scale_sweep.py)."""
    id: str
    family: str
    n: int
    seed: int
    method: str
    certified: bool
    energy: float
    known_energy: float | None
    energy_error: float | None
    wall_s: float


def bank_model(spec: dict):
    """The model of a bank row {id, family, n, seed, params, fingerprint}; the dense fingerprint
    is checked against the row's (drift/benchmarks/bank.py:113-117)."""
    model, _ = build_model(spec["family"], int(spec["n"]), int(spec["seed"]), spec.get("params", {}))
    fp = spec.get("fingerprint")
    if fp is not None and dense_fingerprint(model) != fp:
        raise ValueError(f"{spec['id']}: instance fingerprint {dense_fingerprint(model)} != {fp}")
    return model


def sweep_row(spec: dict, *, exact_max: int = 18, use_gpu: bool = True,
              n_replicas: int | None = None, n_rounds: int = 400,
              sweeps_per_round: int = 4) -> BankRow:
    """Solve one bank row as the reference's scale sweep does (experiments/scale_sweep.py:186-213):
    ``solve`` with the row's seed, 400 rounds of 4 sweeps by default."""
    model = bank_model(spec)
    t0 = time.perf_counter()
    sol = solve(model, exact_max=exact_max, use_gpu=use_gpu, n_replicas=n_replicas,
                n_rounds=n_rounds, sweeps_per_round=sweeps_per_round, seed=int(spec["seed"]))
    wall = time.perf_counter() - t0
    known = spec.get("known_energy")
    err = None if known is None else float(sol.energy) - float(known)
    return BankRow(id=spec["id"], family=spec["family"], n=int(spec["n"]), seed=int(spec["seed"]),
                   method=sol.method, certified=bool(sol.certified), energy=float(sol.energy),
                   known_energy=known, energy_error=err, wall_s=wall)


def planted_bipartite(n: int, p: float, seed: int) -> np.ndarray:
    """Random bipartite graph on two equal halves, edges only across the split: each pair
    (i < n/2, j >= n/2) in row-major order takes one ``default_rng(seed).random()`` draw and is an
    edge when the draw is < p (experiments/phase14_gpu.py:25-35, the same draws in the same
    order).  Its maximum cut is every edge."""
    rng = np.random.default_rng(seed)
    half = n // 2
    draws = rng.random(half * (n - half)).reshape(half, n - half)
    w = np.zeros((n, n))
    w[:half, half:] = (draws < p).astype(float)
    return w + w.T


def bipartite_check(n: int, *, seed: int = 1, n_replicas: int = 256, n_rounds: int = 500,
                    sweeps_per_round: int = 5, pt_seed: int = 0) -> dict:
    """Max cut of a planted bipartite graph (p = min(0.08, 20/n)) by GPU parallel tempering with
    the reference's settings (experiments/phase14_gpu.py:93-97): {"edges", "cut", "optimal"}."""
    from . import gpu

    w = planted_bipartite(n, p=min(0.08, 20.0 / n), seed=seed)
    edges = int(np.count_nonzero(np.triu(w)))
    g = gpu.parallel_tempering_gpu(maxcut_ising(w), n_replicas=n_replicas, T_min=0.05, T_max=6.0,
                                   n_rounds=n_rounds, sweeps_per_round=sweeps_per_round,
                                   seed=pt_seed, record_history=False)
    cut = int(round(cut_value(w, g.best_s)))
    return {"n": n, "edges": edges, "cut": cut, "optimal": cut == edges, "engine": g.engine}


def crystal_check(L: int, *, n_replicas: int = 256, n_rounds: int = 600,
                  sweeps_per_round: int = 5, seed: int = 0) -> dict:
    """The L x L frustrated crystal through ``solve`` with the reference's settings
    (experiments/crystal_gpu.py:28-31): {"energy_per_spin", "period", "striped", "method"}."""
    sol = solve(crystal_2d(L, L), n_replicas=n_replicas, n_rounds=n_rounds,
                sweeps_per_round=sweeps_per_round, seed=seed)
    grid = np.asarray(sol.s).reshape(L, L)
    return {"L": L, "method": sol.method, "energy_per_spin": float(sol.energy) / (L * L),
            "period": int(column_period(grid)), "striped": bool(is_striped(grid))}
