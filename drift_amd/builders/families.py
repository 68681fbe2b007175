"""Seeded dense families of the reference's instance bank (drift/benchmarks/families.py:63-154),
restated so that each (family, n, seed, params) reproduces the reference's (J, h) bit for bit --
their sha256 fingerprints (drift/benchmarks/bank.py:113-117) are checked against the reference
manifest by the tests.  The generators consume numpy's PCG64 stream in the reference's call order.
"""

from __future__ import annotations

import hashlib

import numpy as np

from ..ising import IsingModel
from .crystal import crystal_2d
from .qubo import maxcut_ising, random_graph

FAMILIES = ("maxcut-er", "pmj-glass", "bipartite-maxcut", "ferro-chain", "crystal", "tfim-chain")


def ferro_chain(n: int, j: float = 1.0) -> IsingModel:
    """Open chain with J = j; ground energy -j (n - 1)."""
    J = np.zeros((n, n))
    i = np.arange(n - 1)
    J[i, i + 1] = J[i + 1, i] = j
    return IsingModel(J=J, h=np.zeros(n))


def pmj_spin_glass(n: int, seed: int, *, p: float = 1.0, field: float = 0.0) -> IsingModel:
    """+/-J couplings on G(n, p): per pair (i < j), one uniform decides the bond and, if present,
    ``rng.choice([-1, 1])`` its sign; then an optional N(0, field^2) field."""
    rng = np.random.default_rng(seed)
    J = np.zeros((n, n))
    signs = np.array([-1.0, 1.0])
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                J[i, j] = J[j, i] = float(rng.choice(signs))
    h = rng.normal(scale=field, size=n) if field else np.zeros(n)
    return IsingModel(J=J, h=h)


def bipartite_graph(n: int, p: float, seed: int) -> np.ndarray:
    """Random bipartite graph between the halves [0, n/2) and [n/2, n); its maximum cut is every
    edge."""
    if n % 2:
        raise ValueError(f"bipartite MaxCut family requires even n, got {n}")
    rng = np.random.default_rng(seed)
    half = n // 2
    hit = rng.random(half * (n - half)).reshape(half, n - half) < p
    W = np.zeros((n, n))
    W[:half, half:] = hit
    return W + W.T


def n_edges(W: np.ndarray) -> int:
    return int(np.count_nonzero(np.triu(np.asarray(W))))


def build_model(family: str, n: int, seed: int, params: dict):
    """(model, extra) for one bank row (drift/benchmarks/families.py:119-154)."""
    if family == "maxcut-er":
        W = random_graph(n, p=float(params.get("p", 0.5)), seed=seed)
        return maxcut_ising(W), {"W": W}
    if family == "pmj-glass":
        return pmj_spin_glass(n, seed, p=float(params.get("p", 1.0)),
                              field=float(params.get("field", 0.0))), {}
    if family == "bipartite-maxcut":
        W = bipartite_graph(n, float(params.get("p", 0.5)), seed)
        return maxcut_ising(W), {"W": W, "n_edges": n_edges(W)}
    if family in ("ferro-chain", "tfim-chain"):
        return ferro_chain(n, j=float(params.get("j", 1.0))), {}
    if family == "crystal":
        rows, cols = int(params.get("rows", 4)), int(params.get("cols", 4))
        if rows * cols != n:
            raise ValueError(f"crystal n={n} != rows*cols={rows * cols}")
        return crystal_2d(rows, cols), {}
    raise ValueError(f"unknown instance family {family!r}; expected one of {FAMILIES}")


def analytic_energy(family: str, n: int, extra: dict):
    """Closed-form ground energy where the family has one, else None."""
    if family == "ferro-chain":
        return -float(n - 1)
    if family == "bipartite-maxcut":
        return -float(extra["n_edges"])
    if family == "crystal":
        return -2.0 * n
    return None


def dense_fingerprint(model: IsingModel) -> str:
    """sha256(J || h)[:16] over C-contiguous float64, the bank's convention."""
    payload = np.ascontiguousarray(model.J, dtype=np.float64).tobytes()
    payload += np.ascontiguousarray(model.h, dtype=np.float64).tobytes()
    return hashlib.sha256(payload).hexdigest()[:16]
