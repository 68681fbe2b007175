"""Sparse instance generators for the large parallel-tempering workloads.

All return ``SparseIsingModel`` (CSR with both directions of every bond, columns ascending per
row) and are deterministic in their arguments; ``model.fingerprint()`` pins their output.

* ``square_ferromagnet(L)``        2-D periodic L x L, J = +1, h = 0.
* ``cubic_pmj_glass(L, seed)``     3-D periodic L^3 Edwards-Anderson glass, J = +/-1, h = 0.
* ``er_sparse_weighted(n, d, seed)`` Erdos-Renyi G(n, p = d/(n-1)) with integer weights drawn
                                    uniformly from {-3,-2,-1,1,2,3}, generated in O(n + m) by
                                    geometric skips over the pair sequence (Batagelj & Brandes,
                                    Phys. Rev. E 71, 036113, 2005).
* ``random_regular_gaussian(n, seed)`` random 3-regular graph (pairing model with rejection),
                                    J ~ N(0, 1) and h ~ N(0, 0.1^2), both rounded to float32.

Sites of the hypercubic lattices are numbered with x fastest: i = (z L + y) L + x.
"""

from __future__ import annotations

import numpy as np

from ..ising import SparseIsingModel


def _csr_from_bonds(n: int, a: np.ndarray, b: np.ndarray, w: np.ndarray, h) -> SparseIsingModel:
    """CSR from undirected bonds (a_k, b_k, w_k): both directions, parallel bonds summed, columns
    ascending within each row."""
    rows = np.concatenate([a, b]).astype(np.int64)
    cols = np.concatenate([b, a]).astype(np.int64)
    vals = np.concatenate([w, w]).astype(np.float64)
    key = rows * n + cols
    uniq, inv = np.unique(key, return_inverse=True)
    summed = np.bincount(inv, weights=vals, minlength=uniq.size)
    keep = summed != 0.0
    uniq, summed = uniq[keep], summed[keep]
    r, c = np.divmod(uniq, n)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum(np.bincount(r, minlength=n))
    hv = np.zeros(n) if h is None else np.asarray(h, dtype=np.float64)
    return SparseIsingModel(row_ptr, c, summed, hv, check=False)


def _lattice_bonds(L: int, dim: int):
    """(site, neighbour in +direction d) for d = 0..dim-1, periodic, x fastest."""
    n = L ** dim
    idx = np.arange(n, dtype=np.int64)
    coords = [(idx // L ** d) % L for d in range(dim)]
    out = []
    for d in range(dim):
        nb = idx + (((coords[d] + 1) % L) - coords[d]) * L ** d
        out.append(nb)
    return idx, out


def square_ferromagnet(L: int, *, J: float = 1.0) -> SparseIsingModel:
    """Periodic L x L ferromagnet (config C1 of the parallel-tempering workloads)."""
    idx, nbs = _lattice_bonds(L, 2)
    a = np.concatenate([idx, idx])
    b = np.concatenate(nbs)
    return _csr_from_bonds(L * L, a, b, np.full(a.size, float(J)), None)


def cubic_pmj_glass(L: int, seed: int) -> SparseIsingModel:
    """Periodic L^3 Edwards-Anderson +/-J glass.  Bond (i, i + e_d) takes
    2 * rng.integers(0, 2, (n, 3))[i, d] - 1 with rng = default_rng(seed)."""
    n = L ** 3
    rng = np.random.default_rng(seed)
    signs = (2 * rng.integers(0, 2, size=(n, 3)) - 1).astype(np.float64)
    idx, nbs = _lattice_bonds(L, 3)
    a = np.concatenate([idx, idx, idx])
    b = np.concatenate(nbs)
    w = np.concatenate([signs[:, 0], signs[:, 1], signs[:, 2]])
    return _csr_from_bonds(n, a, b, w, None)


def _pair_from_linear(k: np.ndarray, n: int):
    """Map linear indices over pairs (i < j) in row-major order to (i, j)."""
    # row i starts at off(i) = i (2n - i - 1) / 2
    kf = k.astype(np.float64)
    i = np.floor(((2 * n - 1) - np.sqrt((2 * n - 1) ** 2 - 8 * kf)) / 2).astype(np.int64)
    off = i * (2 * n - i - 1) // 2
    low = off > k
    i[low] -= 1
    off = i * (2 * n - i - 1) // 2
    nxt = (i + 1) * (2 * n - i - 2) // 2
    high = nxt <= k
    i[high] += 1
    off = i * (2 * n - i - 1) // 2
    j = k - off + i + 1
    return i, j


def er_sparse_weighted(n: int, mean_degree: float, seed: int,
                       weights=(-3, -2, -1, 1, 2, 3)) -> SparseIsingModel:
    """Erdos-Renyi G(n, p) with p = mean_degree / (n - 1); each edge takes a weight drawn
    uniformly from ``weights`` (config C3 uses the default integers)."""
    rng = np.random.default_rng(seed)
    total = n * (n - 1) // 2
    p = float(mean_degree) / (n - 1)
    picks = []
    pos = -1
    chunk = max(1024, int(1.2 * p * total) + 1024)
    while True:
        gaps = rng.geometric(p, size=chunk)
        cum = pos + np.cumsum(gaps)
        picks.append(cum[cum < total])
        if cum[-1] >= total:
            break
        pos = int(cum[-1])
    k = np.concatenate(picks)
    i, j = _pair_from_linear(k, n)
    w = np.asarray(weights, dtype=np.float64)[rng.integers(0, len(weights), size=k.size)]
    return _csr_from_bonds(n, i, j, w, None)


def random_regular_gaussian(n: int, seed: int, *, degree: int = 3, field_sigma: float = 0.1,
                            max_tries: int = 10000) -> SparseIsingModel:
    """Random ``degree``-regular graph on n spins by the pairing model (stub permutation,
    rejected until simple), J ~ N(0,1) per edge and h ~ N(0, field_sigma^2), rounded to float32
    so the float engine sees exactly these values."""
    if (n * degree) % 2:
        raise ValueError("n * degree must be even")
    rng = np.random.default_rng(seed)
    stubs = np.repeat(np.arange(n, dtype=np.int64), degree)
    for _ in range(max_tries):
        pairs = stubs[rng.permutation(stubs.size)].reshape(-1, 2)
        a, b = pairs[:, 0], pairs[:, 1]
        if np.any(a == b):
            continue
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        if np.unique(lo * n + hi).size != lo.size:
            continue
        break
    else:  # pragma: no cover - vanishingly unlikely for degree 3
        raise RuntimeError("pairing model did not produce a simple graph")
    J = rng.standard_normal(lo.size).astype(np.float32).astype(np.float64)
    h = (field_sigma * rng.standard_normal(n)).astype(np.float32).astype(np.float64)
    return _csr_from_bonds(n, lo, hi, J, h)
