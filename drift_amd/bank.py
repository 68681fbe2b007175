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
           "crystal_check", "V2_FAMILIES", "V2_GRID", "V2_RUN", "v2_id", "v2_model", "v2_specs",
           "v2_rounds", "v2_best_known_run"]


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


# ------------------------------------------------------------------------------------------------
# bank v2: the families of the parallel-tempering workloads, fingerprinted over their CSR
# ------------------------------------------------------------------------------------------------
V2_FAMILIES = ("ea3d", "er-sparse", "rrg3")

# (family, size parameter, seeds): 3-D +/-J glasses by side L, sparse ER integer graphs (mean
# degree 6, weights +-1..3) and 3-regular N(0,1) graphs (field N(0, 0.1^2)) by spin count
V2_GRID = (
    ("ea3d", (8, 12, 16, 32), (1, 2)),
    ("er-sparse", (1000, 10000, 100000), (1, 2)),
    ("rrg3", (512, 2048, 8192), (1, 2)),
)


def v2_id(family: str, size: int, seed: int) -> str:
    key = "L" if family == "ea3d" else "n"
    return f"v2.{family}.{key}{size:06d}.s{seed:03d}"


def v2_model(family: str, size: int, seed: int):
    """The instance of a v2 row (SparseIsingModel)."""
    from .builders.lattice import cubic_pmj_glass, er_sparse_weighted, random_regular_gaussian

    if family == "ea3d":
        return cubic_pmj_glass(size, seed)
    if family == "er-sparse":
        return er_sparse_weighted(size, 6.0, seed)
    if family == "rrg3":
        return random_regular_gaussian(size, seed)
    raise ValueError(f"unknown v2 family {family!r}; expected one of {V2_FAMILIES}")


def v2_specs() -> list[dict]:
    """Every v2 row: {id, family, size, seed, n, nnz, fingerprint} (CSR sha256[:16], the bank's
    fingerprint convention of drift/benchmarks/bank.py:113-117 applied to the sparse arrays)."""
    from .ising import csr_fingerprint

    out = []
    for family, sizes, seeds in V2_GRID:
        for size in sizes:
            for seed in seeds:
                m = v2_model(family, size, seed)
                out.append({"id": v2_id(family, size, seed), "family": family, "size": size,
                            "seed": seed, "n": m.n, "nnz": m.nnz,
                            "fingerprint": csr_fingerprint(m.row_ptr, m.col_idx, m.weight, m.h)})
    return out


# The best-known runs of every v2 row: PT with R replicas over the geometric ladder T_MIN..T_MAX,
# one sweep and one exchange per round, from each of the independent seeds V2_SEEDS; the row's
# best-known energy is the lowest of them (tests/golden/make_bank_v2.py).  The runs are
# deterministic under the Philox stream contract.
V2_SEEDS = (101, 102, 103, 104)
V2_RUN = {"ea3d": dict(R=64, T_min=1 / 3.0, T_max=5.0),
          "er-sparse": dict(R=32, T_min=0.2, T_max=20.0),
          "rrg3": dict(R=32, T_min=0.1, T_max=3.0)}


def v2_rounds(spec: dict) -> int:
    """Rounds of the recorded best-known run: about 2e11 attempts, at most 200000 rounds."""
    return int(min(200_000, 2e11 // (V2_RUN[spec["family"]]["R"] * spec["n"])))


def v2_best_known_run(spec: dict, n_rounds: int, seed: int = V2_SEEDS[0]):
    """One best-known run of a v2 row (the row's ladder, ``n_rounds`` rounds); returns the
    PtResult."""
    from . import gpu

    cfg = V2_RUN[spec["family"]]
    model = v2_model(spec["family"], spec["size"], spec["seed"])
    return gpu.parallel_tempering_gpu(model, n_replicas=cfg["R"], T_min=cfg["T_min"],
                                      T_max=cfg["T_max"], n_rounds=n_rounds, sweeps_per_round=1,
                                      seed=seed, record_history=False)
