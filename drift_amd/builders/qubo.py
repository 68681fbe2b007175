"""QUBO and MaxCut as Ising models (reference: drift/builders/qubo.py:25-64).

x = (s + 1)/2 turns min x^T Q x into an Ising model with J_ij = -Q_ij/2 (i != j) and
h_i = -1/2 sum_j Q_ij; MaxCut on weights W is the antiferromagnet J = -W, h = 0.
"""

from __future__ import annotations

import numpy as np

from ..ising import IsingModel


def qubo_to_ising(Q: np.ndarray) -> tuple[IsingModel, float]:
    """(model, offset) with E_ising(s) = x^T Q x - offset for x = (s+1)/2."""
    Qs = np.asarray(Q, dtype=np.float64)
    Qs = 0.5 * (Qs + Qs.T)
    J = -0.5 * Qs
    np.fill_diagonal(J, 0.0)
    model = IsingModel(J, -0.5 * Qs.sum(axis=1))
    ones = np.ones(Qs.shape[0])
    return model, float(ones @ Qs @ ones - model.energy(ones))


def maxcut_ising(W: np.ndarray) -> IsingModel:
    """Antiferromagnetic Ising model whose ground state is the maximum cut of W."""
    W = np.asarray(W, dtype=np.float64)
    W = 0.5 * (W + W.T)
    np.fill_diagonal(W, 0.0)
    return IsingModel(-W, np.zeros(W.shape[0]))


def cut_value(W: np.ndarray, s: np.ndarray) -> float:
    """Weight of the edges cut by the partition s: sum_{i<j} W_ij (1 - s_i s_j)/2."""
    W = np.asarray(W, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    return float(0.25 * np.sum(W * (1.0 - np.outer(s, s))))


def random_graph(n: int, p: float = 0.5, seed: int = 0, weighted: bool = False) -> np.ndarray:
    """Erdos-Renyi G(n, p) as a symmetric weight matrix, bit-identical to the reference's.

    The reference draws one uniform per pair (i < j) in row-major order and, for weighted graphs,
    a U(0.5, 2) weight right after each accepted pair.  Unweighted draws are independent of the
    outcomes, so one vectorised call consumes the generator identically."""
    rng = np.random.default_rng(seed)
    W = np.zeros((n, n))
    iu, ju = np.triu_indices(n, k=1)
    if not weighted:
        hit = rng.random(iu.size) < p
        W[iu[hit], ju[hit]] = 1.0
    else:
        for i, j in zip(iu.tolist(), ju.tolist()):
            if rng.random() < p:
                W[i, j] = rng.uniform(0.5, 2.0)
    W = W + W.T
    return W
