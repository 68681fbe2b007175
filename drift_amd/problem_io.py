"""Problem files: a documented sparse format, and the reference's binary wire format.

**drift-csr-v1 (``.npz``).**  The format for instances too large for a dense matrix (N = 10^5,
65536-spin held-out ladders).  One uncompressed or compressed NumPy ``.npz`` archive, readable with
``allow_pickle=False``:

=============  ==========  =====================================================================
key            dtype       content
=============  ==========  =====================================================================
``row_ptr``    int32       n + 1 row offsets of the symmetric coupling matrix, zero diagonal
``col_idx``    int32       nnz column indices, ascending within each row
``weight``     float64     nnz couplings J_ij (both (i, j) and (j, i) stored)
``h``          float64     n fields
``meta``       uint8       UTF-8 JSON: {"format": "drift-csr-v1", "n", "nnz", "fingerprint", ...}
=============  ==========  =====================================================================

Energy convention E(s) = -1/2 s^T J s - h^T s (drift/ising.py:50-60).  ``fingerprint`` is the
sha256[:16] of the CSR arrays (``drift_amd.ising.csr_fingerprint``).

**Reference wire format.**  What ``drift/gpu.py:83-95`` writes and ``cuda/ising_pt.cu:40-45,
159-172`` reads: a packed little-endian header ``<Q i i i i i i f f>`` (seed, n, R, n_rounds,
sweeps_per_round, nnz, k, T_min, T_max), then int32 row_ptr[n+1], int32 col_idx[nnz], float32
weight[nnz], float32 h[n], int32 color_ptr[k+1], int32 color_spins[n].  The result file is a
float64 best energy followed by float32 best_s[n] (``cuda/ising_pt.cu:231-234``,
``drift/gpu.py:101-103``).  ``run_problem_bin`` reads such a problem, runs it on this engine with
the file's parameters and colouring, and writes such a result: the subprocess round trip of the
reference (``drift/gpu.py:97``) without the executable.
"""

from __future__ import annotations

import json
import struct

import numpy as np

from .ising import SparseIsingModel, csr_fingerprint

__all__ = ["save_problem", "load_problem", "write_problem_bin", "read_problem_bin",
           "write_result_bin", "read_result_bin", "run_problem_bin"]

FORMAT = "drift-csr-v1"
_HEADER = struct.Struct("<Qiiiiiiff")


def _csr(model):
    rp, ci, w = model.to_csr()
    return (np.ascontiguousarray(rp, dtype=np.int32), np.ascontiguousarray(ci, dtype=np.int32),
            np.ascontiguousarray(w, dtype=np.float64),
            np.ascontiguousarray(model.h, dtype=np.float64))


def save_problem(path: str, model, *, meta: dict | None = None, compress: bool = True) -> dict:
    """Write ``model`` (dense ``IsingModel`` or ``SparseIsingModel``) as drift-csr-v1; returns
    the metadata written."""
    rp, ci, w, h = _csr(model)
    info = {"format": FORMAT, "n": int(rp.size - 1), "nnz": int(ci.size),
            "fingerprint": csr_fingerprint(rp, ci, w, h)}
    info.update(meta or {})
    blob = np.frombuffer(json.dumps(info, sort_keys=True).encode(), dtype=np.uint8)
    (np.savez_compressed if compress else np.savez)(path, row_ptr=rp, col_idx=ci, weight=w, h=h,
                                                    meta=blob)
    return info


def load_problem(path: str) -> SparseIsingModel:
    """Read a drift-csr-v1 file.  The CSR is validated (shape, ranges, symmetry, zero diagonal)
    by ``SparseIsingModel`` and checked against the stored fingerprint; ``ValueError`` if not."""
    with np.load(path, allow_pickle=False) as z:
        missing = {"row_ptr", "col_idx", "weight", "h", "meta"} - set(z.files)
        if missing:
            raise ValueError(f"{path}: not a {FORMAT} file (missing {sorted(missing)})")
        info = json.loads(bytes(z["meta"]).decode())
        if info.get("format") != FORMAT:
            raise ValueError(f"{path}: format {info.get('format')!r}, expected {FORMAT!r}")
        rp, ci, w, h = (z["row_ptr"].astype(np.int32), z["col_idx"].astype(np.int32),
                        z["weight"].astype(np.float64), z["h"].astype(np.float64))
    model = SparseIsingModel(rp, ci, w, h)
    fp = csr_fingerprint(model.row_ptr, model.col_idx, model.weight, model.h)
    if info.get("fingerprint") not in (None, fp):
        raise ValueError(f"{path}: fingerprint {fp} does not match the stored {info['fingerprint']}")
    model.meta = info
    return model


def write_problem_bin(path: str, model, *, n_replicas: int = 16, n_rounds: int = 300,
                      sweeps_per_round: int = 5, seed: int = 0, T_min: float = 0.1,
                      T_max: float = 5.0, coloring=None) -> None:
    """Write the reference's problem.bin for ``model`` (float32 couplings and field, greedy
    colouring unless one is given), byte for byte what drift/gpu.py:83-95 writes."""
    from . import _lib

    rp, ci, w, h = _csr(model)
    n, nnz = rp.size - 1, ci.size
    if coloring is None:
        cptr = np.zeros(n + 1, dtype=np.int32)
        cspins = np.zeros(max(n, 1), dtype=np.int32)
        lib = _lib.load()
        k = lib.drift_pt_greedy_coloring(n, _lib.ptr(rp, _lib.c_int32), _lib.ptr(ci, _lib.c_int32),
                                         _lib.ptr(cptr, _lib.c_int32), _lib.ptr(cspins, _lib.c_int32))
        if k < 1:
            raise RuntimeError(f"greedy colouring failed ({k}): {_lib.last_error(lib)}")
        cptr, cspins = cptr[: k + 1], cspins[:n]
    else:
        cptr, cspins = (np.ascontiguousarray(c, dtype=np.int32) for c in coloring)
        k = cptr.size - 1
    with open(path, "wb") as f:
        f.write(_HEADER.pack(seed & 0xFFFFFFFFFFFFFFFF, n, n_replicas, n_rounds, sweeps_per_round,
                             nnz, k, float(T_min), float(T_max)))
        for a, dt in ((rp, np.int32), (ci, np.int32), (w, np.float32), (h, np.float32),
                      (cptr, np.int32), (cspins, np.int32)):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def read_problem_bin(path: str) -> dict:
    """Parse a reference problem.bin into its header fields and arrays."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < _HEADER.size:
        raise ValueError(f"{path}: shorter than the {_HEADER.size}-byte header")
    seed, n, R, n_rounds, sweeps, nnz, k, T_min, T_max = _HEADER.unpack_from(buf, 0)
    sizes = [("row_ptr", np.int32, n + 1), ("col_idx", np.int32, nnz), ("weight", np.float32, nnz),
             ("h", np.float32, n), ("color_ptr", np.int32, k + 1), ("color_spins", np.int32, n)]
    need = _HEADER.size + sum(np.dtype(dt).itemsize * c for _, dt, c in sizes)
    if n < 0 or nnz < 0 or k < 1 or len(buf) != need:
        raise ValueError(f"{path}: {len(buf)} bytes, header (n={n}, nnz={nnz}, k={k}) needs {need}")
    out = {"seed": seed, "n": n, "n_replicas": R, "n_rounds": n_rounds, "sweeps_per_round": sweeps,
           "nnz": nnz, "k": k, "T_min": T_min, "T_max": T_max}
    off = _HEADER.size
    for name, dt, count in sizes:
        out[name] = np.frombuffer(buf, dtype=dt, count=count, offset=off).copy()
        off += np.dtype(dt).itemsize * count
    return out


def write_result_bin(path: str, best_E: float, best_s) -> None:
    """The reference's result.bin: float64 best energy, float32 spins."""
    with open(path, "wb") as f:
        f.write(struct.pack("<d", float(best_E)))
        f.write(np.ascontiguousarray(best_s, dtype=np.float32).tobytes())


def read_result_bin(path: str, n: int):
    """(best_E, best_s as float64) from a result.bin of n spins (drift/gpu.py:101-103)."""
    with open(path, "rb") as f:
        best_E = struct.unpack("<d", f.read(8))[0]
        best_s = np.frombuffer(f.read(4 * n), dtype=np.float32).astype(np.float64).copy()
    return float(best_E), best_s


def run_problem_bin(problem_path: str, result_path: str, **kw):
    """Solve a reference problem.bin on this engine with the file's seed, replica count, rounds,
    sweeps per round, temperature range and colouring; write result.bin; return the PtResult.
    Keyword arguments go to ``parallel_tempering_gpu`` (engine, ladder, ...)."""
    from . import gpu

    p = read_problem_bin(problem_path)
    model = SparseIsingModel(p["row_ptr"], p["col_idx"], p["weight"].astype(np.float64),
                             p["h"].astype(np.float64))
    res = gpu.parallel_tempering_gpu(
        model, n_replicas=p["n_replicas"], T_min=float(p["T_min"]), T_max=float(p["T_max"]),
        n_rounds=p["n_rounds"], sweeps_per_round=p["sweeps_per_round"], seed=p["seed"],
        coloring=(p["color_ptr"], p["color_spins"]), **kw)
    write_result_bin(result_path, res.best_E, res.best_s)
    return res
