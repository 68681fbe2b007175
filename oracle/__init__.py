"""CPU oracle of the parallel-tempering path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of liboracle_pt.so (pt_oracle.c): Philox4x32-10, the deterministic exponentials,
greedy colouring, the colour-ordered parallel-tempering chain under the engine's stream contract,
and Gray-code exact enumeration with thermal averages.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg import this package; the product (drift_amd/) never does.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int8, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pt.so")
FAST_LIB_PATH = os.path.join(HERE, "liboracle_pt_fast.so")   # -O3 build for CPU-baseline timing

W_I8, W_I32, W_F32 = 1, 2, 3


class OrProblem(ctypes.Structure):
    _fields_ = [
        ("n", c_int32), ("nnz", c_int32),
        ("row_ptr", POINTER(c_int32)), ("col_idx", POINTER(c_int32)),
        ("weight_type", c_int32),
        ("weight", c_void_p), ("h", c_void_p),
        ("energy_scale", c_double),
        ("n_colors", c_int32),
        ("color_ptr", POINTER(c_int32)), ("color_spins", POINTER(c_int32)),
    ]


class OrParams(ctypes.Structure):
    _fields_ = [
        ("seed", c_uint64), ("R", c_int32), ("n_rounds", c_int32),
        ("sweeps_per_round", c_int32), ("n_threads", c_int32),
        ("beta", POINTER(c_double)),
    ]


class OrResult(ctypes.Structure):
    _fields_ = [
        ("best_s", POINTER(c_int8)),
        ("best_units", c_int64),
        ("best_E", c_double),
        ("best_round", c_int32), ("best_rung", c_int32), ("best_slot", c_int32), ("pad", c_int32),
        ("swap_tried", POINTER(c_int64)), ("swap_acc", POINTER(c_int64)),
        ("round_trips", POINTER(c_int32)),
        ("rung_up", POINTER(c_int64)), ("rung_down", POINTER(c_int64)),
        ("rung_of", POINTER(c_int32)),
        ("final_units", POINTER(c_int64)),
        ("final_spins", POINTER(c_int8)),
        ("best_history", POINTER(c_double)),
        ("rung_trace", POINTER(c_double)),
        ("seconds", c_double),
    ]


_lib = None
_fast = None


def build() -> None:
    """Compile liboracle_pt.so from pt_oracle.c (make, gcc)."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def load(fast: bool = False):
    """The checker build, or (fast) the -O3 build of the same source used to time the CPU
    baseline; both run the same chain bit for bit."""
    global _lib, _fast
    if fast:
        if _fast is None:
            if not os.path.exists(FAST_LIB_PATH):
                build()
            _fast = _bind(ctypes.CDLL(FAST_LIB_PATH))
        return _fast
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def _bind(lib):
    """Argument and result types of the oracle's exports."""
    lib.or_philox.argtypes = [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)]
    lib.or_philox.restype = None
    lib.or_exp.argtypes = [c_double]
    lib.or_exp.restype = c_double
    lib.or_expf.argtypes = [c_float]
    lib.or_expf.restype = c_float
    lib.or_thr_int.argtypes = [c_double, c_int32]
    lib.or_thr_int.restype = c_uint32
    lib.or_greedy_coloring.argtypes = [c_int32] + [POINTER(c_int32)] * 4
    lib.or_greedy_coloring.restype = c_int32
    lib.or_pt_run.argtypes = [POINTER(OrProblem), POINTER(OrParams), POINTER(OrResult)]
    lib.or_pt_run.restype = c_int32
    lib.or_exact.argtypes = [c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_double),
                             POINTER(c_double), c_int32, POINTER(c_double), POINTER(c_double),
                             POINTER(c_int64), POINTER(c_double), POINTER(c_double)]
    lib.or_exact.restype = c_int32
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(POINTER(t))


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().or_philox(_p(c, c_uint32), _p(k, c_uint32), _p(out, c_uint32))
    return out


def exp_det(x: float) -> float:
    return load().or_exp(float(x))


def expf_det(x: float) -> float:
    return load().or_expf(float(x))


def thr_int(beta_scaled: float, q: int) -> int:
    return int(load().or_thr_int(float(beta_scaled), int(q)))


def greedy_coloring(row_ptr, col_idx):
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    n = rp.size - 1
    cptr = np.zeros(n + 2, dtype=np.int32)
    cs = np.zeros(max(n, 1), dtype=np.int32)
    k = load().or_greedy_coloring(n, _p(rp, c_int32), _p(ci, c_int32), _p(cptr, c_int32), _p(cs, c_int32))
    return int(k), cptr[: k + 1].copy(), cs[:n].copy()


def pt_run(n, row_ptr, col_idx, weight_type, weight, h, energy_scale, color_ptr, color_spins,
           *, seed, R, n_rounds, sweeps_per_round, beta, n_threads=0, trace=False, fast=False):
    """Run the CPU chain; returns a dict of numpy arrays and scalars (fast: the -O3 build)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    cp = np.ascontiguousarray(color_ptr, dtype=np.int32)
    cs = np.ascontiguousarray(color_spins, dtype=np.int32)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    w = np.ascontiguousarray(weight)
    hh = None if h is None else np.ascontiguousarray(h)
    prob = OrProblem(n=n, nnz=int(ci.size), row_ptr=_p(rp, c_int32), col_idx=_p(ci, c_int32),
                     weight_type=weight_type, weight=ctypes.c_void_p(w.ctypes.data),
                     h=None if hh is None else ctypes.c_void_p(hh.ctypes.data),
                     energy_scale=float(energy_scale), n_colors=int(cp.size - 1),
                     color_ptr=_p(cp, c_int32), color_spins=_p(cs, c_int32))
    prm = OrParams(seed=int(seed) & 0xFFFFFFFFFFFFFFFF, R=R, n_rounds=n_rounds,
                   sweeps_per_round=sweeps_per_round, n_threads=n_threads, beta=_p(b, c_double))
    out = {
        "best_s": np.zeros(n, dtype=np.int8),
        "swap_tried": np.zeros(max(R - 1, 1), dtype=np.int64),
        "swap_acc": np.zeros(max(R - 1, 1), dtype=np.int64),
        "round_trips": np.zeros(R, dtype=np.int32),
        "rung_up": np.zeros(R, dtype=np.int64),
        "rung_down": np.zeros(R, dtype=np.int64),
        "rung_of": np.zeros(R, dtype=np.int32),
        "final_units": np.zeros(R, dtype=np.int64),
        "final_spins": np.zeros((R, n), dtype=np.int8),
        "best_history": np.zeros(max(n_rounds, 1), dtype=np.float64),
        "rung_trace": np.zeros((max(n_rounds, 1), R), dtype=np.float64) if trace else None,
    }
    res = OrResult(
        best_s=_p(out["best_s"], c_int8), swap_tried=_p(out["swap_tried"], c_int64),
        swap_acc=_p(out["swap_acc"], c_int64), round_trips=_p(out["round_trips"], c_int32),
        rung_up=_p(out["rung_up"], c_int64), rung_down=_p(out["rung_down"], c_int64),
        rung_of=_p(out["rung_of"], c_int32), final_units=_p(out["final_units"], c_int64),
        final_spins=_p(out["final_spins"], c_int8), best_history=_p(out["best_history"], c_double),
        rung_trace=_p(out["rung_trace"], c_double))
    rc = load(fast).or_pt_run(prob, prm, res)
    if rc != 0:
        raise RuntimeError(f"oracle rejected the problem ({rc})")
    out["swap_tried"] = out["swap_tried"][: max(R - 1, 0)]
    out["swap_acc"] = out["swap_acc"][: max(R - 1, 0)]
    out["best_history"] = out["best_history"][:n_rounds]
    if trace:
        out["rung_trace"] = out["rung_trace"][:n_rounds]
    out.update(best_units=int(res.best_units), best_E=float(res.best_E),
               best_round=int(res.best_round), best_rung=int(res.best_rung),
               best_slot=int(res.best_slot), seconds=float(res.seconds))
    return out


def pt_run_encoded(enc, coloring, **kw):
    """pt_run on a drift_amd.gpu.EncodedProblem and an explicit (color_ptr, color_spins)."""
    cptr, cspins = coloring
    return pt_run(enc.n, enc.row_ptr, enc.col_idx, enc.weight_type, enc.weight, enc.h, enc.scale,
                  cptr, cspins, **kw)


def exact(row_ptr, col_idx, weight, h, betas=()):
    """(e_min, idx_min, mean_E, mean_E2) by Gray-code enumeration (n <= 26)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    w = np.ascontiguousarray(weight, dtype=np.float64)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    b = np.ascontiguousarray(np.atleast_1d(betas), dtype=np.float64)
    nb = int(b.size)
    mE = np.zeros(max(nb, 1))
    mE2 = np.zeros(max(nb, 1))
    emin = c_double()
    imin = c_int64()
    rc = load().or_exact(rp.size - 1, _p(rp, c_int32), _p(ci, c_int32), _p(w, c_double),
                         _p(hh, c_double), nb, _p(b, c_double) if nb else None,
                         ctypes.byref(emin), ctypes.byref(imin), _p(mE, c_double), _p(mE2, c_double))
    if rc != 0:
        raise ValueError("oracle exact enumeration refused the instance")
    return float(emin.value), int(imin.value), mE[:nb], mE2[:nb]
