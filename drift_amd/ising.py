"""Ising models: the dense reference type and a sparse (CSR) type for large instances.

Energy convention (drift/ising.py:1-18 of the reference): for s in {-1, +1}^n,

    E(s) = -1/2 s^T J s - h^T s,    dE(flip i) = 2 s_i (J_i . s + h_i),

with J symmetric and zero on the diagonal.  ``IsingModel`` keeps the reference's dense float64
semantics (drift/ising.py:25-78), including the symmetrisation J <- (J + J^T)/2 and the zeroed
diagonal.  ``SparseIsingModel`` holds the same object as CSR so that instances with 10^5 spins
(where a dense J would need 80 GB) can reach the engine; it offers the same methods.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np


@dataclass
class IsingModel:
    """Dense classical Ising model (reference: drift/ising.py:25-78)."""

    J: np.ndarray
    h: np.ndarray

    def __post_init__(self) -> None:
        self.J = np.asarray(self.J, dtype=np.float64)
        self.h = np.asarray(self.h, dtype=np.float64)
        if self.h.ndim != 1:
            raise ValueError("h must be a 1-D field vector")
        n = self.h.shape[0]
        if self.J.shape != (n, n):
            raise ValueError(f"J must be ({n}, {n}) to match h, got {self.J.shape}")
        self.J = 0.5 * (self.J + self.J.T)
        np.fill_diagonal(self.J, 0.0)

    @property
    def n(self) -> int:
        return self.h.shape[0]

    def energy(self, s: np.ndarray) -> float:
        s = np.asarray(s, dtype=np.float64)
        return float(-0.5 * s @ self.J @ s - self.h @ s)

    def energy_batch(self, S: np.ndarray) -> np.ndarray:
        S = np.asarray(S, dtype=np.float64)
        return -0.5 * np.einsum("mi,mi->m", S @ self.J, S) - S @ self.h

    def delta_energy_flip(self, s: np.ndarray, i: int) -> float:
        return float(2.0 * s[i] * (self.J[i] @ s + self.h[i]))

    def local_fields(self, s: np.ndarray) -> np.ndarray:
        return self.J @ s + self.h

    @property
    def nbonds(self) -> int:
        return int(np.count_nonzero(np.triu(self.J)))

    def to_csr(self):
        """(row_ptr, col_idx, weight) of the nonzero couplings, rows grouped and columns ascending
        (the order np.nonzero gives, as in drift/gpu.py:69-77)."""
        rows, cols = np.nonzero(self.J)
        weight = self.J[rows, cols]
        row_ptr = np.zeros(self.n + 1, dtype=np.int32)
        row_ptr[1:] = np.cumsum(np.bincount(rows, minlength=self.n))
        return row_ptr, cols.astype(np.int32), weight.astype(np.float64)

    @classmethod
    def from_csr(cls, row_ptr, col_idx, weight, h) -> "SparseIsingModel":
        return SparseIsingModel(row_ptr, col_idx, weight, h)


class SparseIsingModel:
    """The same Ising object stored as CSR (both directions of every bond, no diagonal).

    Rows keep the caller's column order; that order is the fixed summation order of the float
    engine.  Duplicate (i, j) entries are merged by summing, as a dense J would.
    """

    def __init__(self, row_ptr, col_idx, weight, h, *, check: bool = True):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        col_idx = np.ascontiguousarray(col_idx, dtype=np.int64)
        weight = np.ascontiguousarray(weight, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        if h.ndim != 1:
            raise ValueError("h must be a 1-D field vector")
        n = h.shape[0]
        if row_ptr.shape != (n + 1,) or row_ptr[0] != 0 or row_ptr[-1] != col_idx.size:
            raise ValueError("row_ptr must have n+1 entries running from 0 to nnz")
        if weight.shape != col_idx.shape:
            raise ValueError("weight and col_idx must have the same length")
        if np.any(np.diff(row_ptr) < 0):
            raise ValueError("row_ptr must be non-decreasing")
        if col_idx.size and (col_idx.min() < 0 or col_idx.max() >= n):
            raise ValueError("col_idx out of range")
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(row_ptr))
        if check:
            keep = rows != col_idx                       # the diagonal carries no energy
            rows, col_idx, weight = rows[keep], col_idx[keep], weight[keep]
            key = rows * n + col_idx
            uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
            if uniq.size != key.size:                    # merge duplicates, first occurrence order
                summed = np.bincount(inv, weights=weight, minlength=uniq.size)
                order = np.sort(first)
                rows, col_idx = rows[order], col_idx[order]
                weight = summed[inv[order]]
                nz = weight != 0.0
                rows, col_idx, weight = rows[nz], col_idx[nz], weight[nz]
                key = rows * n + col_idx
            tkey = col_idx * n + rows
            a, b = np.argsort(key), np.argsort(tkey)
            if not (np.array_equal(key[a], tkey[b]) and np.array_equal(weight[a], weight[b])):
                raise ValueError("couplings must be symmetric (store both directions of each bond)")
            row_ptr = np.zeros(n + 1, dtype=np.int64)
            row_ptr[1:] = np.cumsum(np.bincount(rows, minlength=n))
        self.row_ptr = row_ptr.astype(np.int32)
        self.col_idx = col_idx.astype(np.int32)
        self.weight = weight
        self.h = h
        self._rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(self.row_ptr))

    @property
    def n(self) -> int:
        return self.h.shape[0]

    @property
    def nnz(self) -> int:
        return int(self.col_idx.size)

    @property
    def nbonds(self) -> int:
        return self.nnz // 2

    def to_csr(self):
        return self.row_ptr, self.col_idx, self.weight

    def local_fields(self, s: np.ndarray) -> np.ndarray:
        s = np.asarray(s, dtype=np.float64)
        lj = np.bincount(self._rows, weights=self.weight * s[self.col_idx], minlength=self.n)
        return lj + self.h

    def energy(self, s: np.ndarray) -> float:
        s = np.asarray(s, dtype=np.float64)
        lj = np.bincount(self._rows, weights=self.weight * s[self.col_idx], minlength=self.n)
        return float(-0.5 * s @ lj - self.h @ s)

    def energy_batch(self, S: np.ndarray) -> np.ndarray:
        S = np.atleast_2d(np.asarray(S, dtype=np.float64))
        return np.array([self.energy(s) for s in S])

    def delta_energy_flip(self, s: np.ndarray, i: int) -> float:
        lo, hi = self.row_ptr[i], self.row_ptr[i + 1]
        return float(2.0 * s[i] * (self.weight[lo:hi] @ np.asarray(s)[self.col_idx[lo:hi]] + self.h[i]))

    def to_dense(self) -> IsingModel:
        J = np.zeros((self.n, self.n))
        J[self._rows, self.col_idx] = self.weight
        return IsingModel(J=J, h=self.h.copy())

    def fingerprint(self) -> str:
        """sha256 of the CSR arrays and h, first 16 hex digits (the reference fingerprints dense
        (J, h) the same way, drift/benchmarks/bank.py:113-117)."""
        return csr_fingerprint(self.row_ptr, self.col_idx, self.weight, self.h)


def csr_fingerprint(row_ptr, col_idx, weight, h) -> str:
    payload = np.ascontiguousarray(row_ptr, dtype="<i4").tobytes()
    payload += np.ascontiguousarray(col_idx, dtype="<i4").tobytes()
    payload += np.ascontiguousarray(weight, dtype="<f8").tobytes()
    payload += np.ascontiguousarray(h, dtype="<f8").tobytes()
    return hashlib.sha256(payload).hexdigest()[:16]


def as_csr(model):
    """(n, row_ptr, col_idx, weight, h) for either model type."""
    if hasattr(model, "to_csr"):
        row_ptr, col_idx, weight = model.to_csr()
        return model.n, row_ptr, col_idx, weight, np.asarray(model.h, dtype=np.float64)
    raise TypeError(f"unsupported model type {type(model).__name__}")
