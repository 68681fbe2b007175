"""Exact enumeration for small n (reference: drift/solvers/exact.py:20-49).

Configuration index k maps to spins by bit b -> spin (1 - 2 b): bit 0 -> +1, bit 1 -> -1.  The
ground state is the first minimum in index order (numpy argmin).  Enumeration runs in chunks of
2^16 configurations so n = 24 needs ~130 MB for the energies rather than 3.2 GB of spins.
``exact_thermal`` adds the exact Boltzmann averages <E>_beta and <E^2>_beta, the reference values
the thermal checks of the parallel-tempering engine are held to.
"""

from __future__ import annotations

import numpy as np

MAX_N = 24
_CHUNK_BITS = 16


def all_configs(n: int) -> np.ndarray:
    """All 2^n configurations as a (2^n, n) float array of +/-1 (row k <-> index k)."""
    if n > MAX_N:
        raise ValueError(f"2^{n} configurations is too many to enumerate")
    idx = np.arange(1 << n, dtype=np.int64)
    bits = (idx[:, None] >> np.arange(n, dtype=np.int64)[None, :]) & 1
    return (1 - 2 * bits).astype(np.float64)


def _dense(model):
    return model if hasattr(model, "J") else model.to_dense()


def _chunks(n: int):
    """Yield (start, S) blocks of configurations in index order."""
    low = min(n, _CHUNK_BITS)
    base = all_configs(low)
    for hi in range(1 << (n - low)):
        if n > low:
            hb = (hi >> np.arange(n - low)) & 1
            top = np.broadcast_to((1 - 2 * hb).astype(np.float64), (base.shape[0], n - low))
            yield hi << low, np.hstack([base, top])
        else:
            yield 0, base


def _energies(model) -> np.ndarray:
    m = _dense(model)
    out = np.empty(1 << m.n, dtype=np.float64)
    for start, S in _chunks(m.n):
        out[start:start + S.shape[0]] = m.energy_batch(S)
    return out


def exact_ground_state(model, max_n: int = 22):
    """(s_min, e_min, all_energies) by brute force; raises for n > max_n."""
    n = model.n
    if n > max_n:
        raise ValueError(
            f"exact ground state only up to {max_n} spins (got {n}); use annealing instead")
    if n > MAX_N:
        raise ValueError(f"2^{n} configurations is too many to enumerate")
    E = _energies(model)
    i = int(np.argmin(E))
    s = (1 - 2 * ((i >> np.arange(n)) & 1)).astype(np.float64)
    return s, float(E[i]), E


def exact_thermal(model, betas):
    """Exact <E> and <E^2> at each inverse temperature, plus the ground energy.

    Returns (mean_E, mean_E2, e_min) with mean_E, mean_E2 shaped like betas."""
    if model.n > MAX_N:
        raise ValueError(f"2^{model.n} configurations is too many to enumerate")
    betas = np.atleast_1d(np.asarray(betas, dtype=np.float64))
    E = _energies(model)
    e0 = float(E.min())
    x = E - e0
    mean_E = np.empty_like(betas)
    mean_E2 = np.empty_like(betas)
    for k, b in enumerate(betas):
        w = np.exp(-b * x)
        Z = w.sum()
        mean_E[k] = float((w * E).sum() / Z)
        mean_E2[k] = float((w * E * E).sum() / Z)
    return mean_E, mean_E2, e0
