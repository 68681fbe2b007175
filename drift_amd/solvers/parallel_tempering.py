"""Replica-exchange Monte Carlo on the CPU, the temperature ladder, and the result record.

``geometric_ladder`` and ``PtResult`` are the shared surface of the CPU and GPU paths
(reference: drift/solvers/parallel_tempering.py:29-49).  ``parallel_tempering`` is the portable
CPU method that ``drift_amd.solve`` uses when no GPU engine is present: random-site Metropolis on
every replica, then proposals to exchange the configurations of adjacent rungs, even pairs on even
rounds and odd pairs on odd rounds (reference: parallel_tempering.py:52-124).  It is a separate,
labelled method ("cpu-pt"); the GPU entry point never falls back to it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def geometric_ladder(n_replicas: int, T_min: float, T_max: float) -> np.ndarray:
    """Ascending temperatures T_k = T_min (T_max/T_min)^(k/(R-1)); one rung gives [T_min]."""
    if n_replicas == 1:
        return np.array([T_min], dtype=np.float64)
    k = np.arange(n_replicas) / (n_replicas - 1)
    return T_min * (T_max / T_min) ** k


def ladder_betas(temperatures: np.ndarray) -> np.ndarray:
    """beta = 1 / max(T, 1e-12), the reference's guard (parallel_tempering.py:72-73)."""
    return 1.0 / np.maximum(np.asarray(temperatures, dtype=np.float64), 1e-12)


@dataclass
class PtResult:
    """Outcome of a parallel-tempering run.  The first seven fields are the reference's
    (drift/solvers/parallel_tempering.py:39-49); the rest are additive diagnostics the GPU engine
    fills (None elsewhere)."""

    best_s: np.ndarray
    best_E: float
    temperatures: np.ndarray
    swap_rate: float
    best_history: np.ndarray = field(default=None)
    throughput: float = None        # spin-flip attempts per second (sweep loop only)
    seconds: float = None           # device time of the sweep + exchange loop
    swap_acceptance: np.ndarray = None   # per adjacent rung pair
    round_trips: np.ndarray = None       # per replica slot
    rung_up: np.ndarray = None           # per rung, visits by replicas last at the cold end
    rung_down: np.ndarray = None         # per rung, visits by replicas last at the hot end
    rung_of: np.ndarray = None           # final rung of each replica slot
    final_E: np.ndarray = None           # final energy of each replica slot
    final_spins: np.ndarray = None       # (R, n) final configurations by slot
    best_round: int = None               # 0 = initial state, r+1 = after round r
    best_rung: int = None
    best_E_units: int = None             # exact integer (or 2^-36 fixed-point) energy
    engine: str = None
    rung_energy_trace: np.ndarray = None # (n_rounds, R) energy at each rung after the sweeps
    swap_tried: np.ndarray = None        # per adjacent rung pair
    swap_accepted: np.ndarray = None
    kernel_seconds: float = None         # device time of the timed sweep-kernel launches
    kernel_attempts: float = None        # attempts those launches performed
    kernel_launches: int = None
    other_kernel_seconds: float = None
    other_kernel_launches: int = None
    rounds_run: int = None               # rounds run (fewer than asked when a target was reached)
    block_threads: int = None            # workgroup size of the sweep kernel
    slots_per_group: int = None          # replica slots per workgroup (split and LDS engines)
    couplings_in_lds: bool = None        # split engine: couplings staged in LDS
    parts_per_group: int = None          # split engine: workgroups per replica group


def _neighbour_access(model):
    """Return f(i) -> (cols, weights) for the couplings of spin i."""
    if hasattr(model, "J"):
        J = model.J
        return lambda i: (slice(None), J[:, i])
    rp, ci, w = model.row_ptr, model.col_idx, model.weight
    return lambda i: (ci[rp[i]:rp[i + 1]], w[rp[i]:rp[i + 1]])


def parallel_tempering(
    model,
    *,
    n_replicas: int = 16,
    T_min: float = 0.1,
    T_max: float = 5.0,
    n_rounds: int = 300,
    sweeps_per_round: int = 5,
    seed: int = 0,
) -> PtResult:
    """CPU parallel tempering: returns the lowest configuration seen over all replicas and rounds
    (initial states included), with the swap acceptance rate and the per-round best history."""
    rng = np.random.default_rng(seed)
    n, R = model.n, n_replicas
    temps = geometric_ladder(R, T_min, T_max)
    beta = ladder_betas(temps)
    spins = (rng.integers(0, 2, size=(R, n)) * 2 - 1).astype(np.float64)
    energies = np.array([model.energy(s) for s in spins])
    fields = np.array([model.local_fields(s) for s in spins])
    couplings = _neighbour_access(model)

    b0 = int(np.argmin(energies))
    best_s, best_E = spins[b0].copy(), float(energies[b0])
    history = np.empty(n_rounds, dtype=np.float64)
    tried = accepted = 0
    attempts = sweeps_per_round * n
    for rnd in range(n_rounds):
        for r in range(R):
            s, f, e, b = spins[r], fields[r], energies[r], beta[r]
            sites = rng.integers(n, size=attempts)
            draws = rng.random(attempts)
            for i, u in zip(sites.tolist(), draws.tolist()):
                dE = 2.0 * s[i] * f[i]
                if dE <= 0.0 or u < math.exp(-dE * b):
                    s[i] = -s[i]
                    e += dE
                    cols, w = couplings(i)
                    f[cols] += 2.0 * s[i] * w
            energies[r] = e
            if e < best_E:
                best_E, best_s = float(e), s.copy()
        for r in range(rnd % 2, R - 1, 2):
            tried += 1
            d = (beta[r] - beta[r + 1]) * (energies[r] - energies[r + 1])
            if d >= 0.0 or rng.random() < math.exp(d):
                accepted += 1
                spins[[r, r + 1]] = spins[[r + 1, r]]
                fields[[r, r + 1]] = fields[[r + 1, r]]
                energies[[r, r + 1]] = energies[[r + 1, r]]
        history[rnd] = best_E
    return PtResult(
        best_s=best_s,
        best_E=best_E,
        temperatures=temps,
        swap_rate=accepted / tried if tried else 0.0,
        best_history=history,
    )
