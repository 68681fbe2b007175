"""Period-4 stripe crystal (reference: drift/builders/crystal.py:28-80).

Along x a ferromagnetic nearest-neighbour bond j1 and an antiferromagnetic next-nearest bond j2;
along y a ferromagnetic bond jy.  For |j2| > j1/2 the ground state is vertical stripes of width 2,
E = -2n when cols is a multiple of 4 (periodic).
"""

from __future__ import annotations

import numpy as np

from ..ising import IsingModel


def crystal_2d(rows: int, cols: int, *, j1: float = 1.0, j2: float = -1.0, jy: float = 1.0,
               periodic: bool = True) -> IsingModel:
    n = rows * cols
    J = np.zeros((n, n))
    r, c = np.divmod(np.arange(n), cols)
    bonds = []
    for dc, w in ((1, j1), (2, j2)):
        cc = c + dc
        ok = cc < cols if not periodic else np.ones(n, dtype=bool)
        bonds.append((np.arange(n)[ok], (r * cols + cc % cols)[ok], w))
    rr = r + 1
    ok = rr < rows if not periodic else np.ones(n, dtype=bool)
    bonds.append((np.arange(n)[ok], ((rr % rows) * cols + c)[ok], jy))
    for a, b, w in bonds:      # accumulate like the reference: J[a,b] += w; J[b,a] += w
        np.add.at(J, (a, b), w)
        np.add.at(J, (b, a), w)
    return IsingModel(J, np.zeros(n))


def column_period(grid: np.ndarray) -> int:
    """Smallest horizontal period p with column c == column c+p for all c (else cols)."""
    grid = np.asarray(grid)
    cols = grid.shape[1]
    for p in range(1, cols):
        if cols % p == 0 and np.all(grid == np.roll(grid, p, axis=1)):
            return p
    return cols


def is_striped(grid: np.ndarray) -> bool:
    """True for the defect-free period-4 stripe crystal."""
    return column_period(np.asarray(grid)) == 4
