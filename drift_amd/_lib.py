"""ctypes binding of libdriftpt.so (include/drift_pt.h).

This replaces the reference's process boundary -- drift/gpu.py:83-106 wrote problem.bin, ran the
CUDA binary in a subprocess and parsed result.bin and stderr -- with one in-process C call.  The
structures below mirror include/drift_pt.h field for field.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int8, c_int32, c_int64, c_uint64, c_void_p

LIB_NAME = "libdriftpt.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

ABI_VERSION = 6

# enum drift_pt_status
OK, EINVAL, ECOLOR, ENOMEM, EHIP, ENODEV, ERANGE = 0, -1, -2, -3, -4, -5, -6
# enum drift_pt_weight_type
W_I8, W_I32, W_F32 = 1, 2, 3
# enum drift_pt_engine
ENGINE_AUTO, ENGINE_LDS, ENGINE_GLOBAL, ENGINE_LATTICE, ENGINE_SLAB, ENGINE_COOP = 0, 1, 2, 3, 4, 5
ENGINE_PSLAB = 6
ENGINE_SPLIT = 7
ENGINES = {"auto": ENGINE_AUTO, "lds": ENGINE_LDS, "global": ENGINE_GLOBAL,
           "lattice": ENGINE_LATTICE, "slab": ENGINE_SLAB, "coop": ENGINE_COOP,
           "pslab": ENGINE_PSLAB, "split": ENGINE_SPLIT}
ENGINE_NAMES = {v: k for k, v in ENGINES.items()}
# enum drift_pt_flags
RECORD_HISTORY, RECORD_FINAL, SKIP_VALIDATION, RECORD_TRACE, TIME_KERNELS, NO_GRAPH = 1, 2, 4, 8, 16, 32
AUTOTUNE = 64
PLAIN_LAUNCH = 128
STOP_AT_TARGET = 256
COUPLINGS_LDS = 512
TUNE_MAX = 64
NOALLOC_BYTES = 64 << 20   # DRIFT_PT_NOALLOC_BYTES


class Problem(ctypes.Structure):
    _fields_ = [
        ("n", c_int32),
        ("nnz", c_int32),
        ("row_ptr", POINTER(c_int32)),
        ("col_idx", POINTER(c_int32)),
        ("weight_type", c_int32),
        ("n_colors", c_int32),
        ("weight", c_void_p),
        ("h", c_void_p),
        ("energy_scale", c_double),
        ("color_ptr", POINTER(c_int32)),
        ("color_spins", POINTER(c_int32)),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("seed", c_uint64),
        ("n_replicas", c_int32),
        ("n_rounds", c_int32),
        ("sweeps_per_round", c_int32),
        ("engine", c_int32),
        ("T_min", c_double),
        ("T_max", c_double),
        ("beta", POINTER(c_double)),
        ("flags", c_int32),
        ("block_threads", c_int32),
        ("target_energy", c_double),
        ("slots_per_group", c_int32),
        ("parts_per_group", c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("best_s", POINTER(c_int8)),
        ("best_E", c_double),
        ("best_E_units", c_int64),
        ("best_round", c_int32),
        ("best_rung", c_int32),
        ("best_slot", c_int32),
        ("engine_used", c_int32),
        ("swap_tried", POINTER(c_int64)),
        ("swap_accepted", POINTER(c_int64)),
        ("round_trips", POINTER(c_int32)),
        ("rung_up", POINTER(c_int64)),
        ("rung_down", POINTER(c_int64)),
        ("rung_of", POINTER(c_int32)),
        ("final_E_units", POINTER(c_int64)),
        ("final_spins", POINTER(c_int8)),
        ("best_history", POINTER(c_double)),
        ("rung_energy_trace", POINTER(c_double)),
        ("seconds", c_double),
        ("attempts", c_double),
        ("kernel_seconds", c_double),
        ("kernel_attempts", c_double),
        ("kernel_launches", c_int64),
        ("other_kernel_seconds", c_double),
        ("other_kernel_launches", c_int64),
        ("rounds_run", c_int32),
        ("block_threads_used", c_int32),
        ("slots_per_group_used", c_int32),
        ("couplings_in_lds_used", c_int32),
        ("parts_per_group_used", c_int32),
    ]


# every symbol include/drift_pt.h declares, with its signature
class Tuning(ctypes.Structure):
    """drift_pt_tuning: the timed candidates of one autotune call."""
    _fields_ = [
        ("engine", c_int32),
        ("block_threads", c_int32),
        ("slots_per_group", c_int32),
        ("couplings_in_lds", c_int32),
        ("parts_per_group", c_int32),
        ("trial_rounds", c_int32),
        ("n_candidates", c_int32),
        ("cand_engine", c_int32 * TUNE_MAX),
        ("cand_block", c_int32 * TUNE_MAX),
        ("cand_group", c_int32 * TUNE_MAX),
        ("cand_lds_couplings", c_int32 * TUNE_MAX),
        ("cand_parts", c_int32 * TUNE_MAX),
        ("cand_seconds", c_double * TUNE_MAX),
    ]


SIGNATURES = {
    "drift_pt_abi_version": ([], c_int),
    "drift_pt_device_available": ([], c_int),
    "drift_pt_last_error": ([], c_char_p),
    "drift_pt_greedy_coloring": (
        [c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)], c_int),
    "drift_pt_run": ([POINTER(Problem), POINTER(Params), POINTER(Result)], c_int),
    "drift_pt_run_batch": ([POINTER(Problem), POINTER(Params), c_int32, POINTER(Result)], c_int),
    "drift_pt_run_batch_noalloc": ([POINTER(Problem), POINTER(Params), c_int32, POINTER(Result)],
                                   c_int),
    "drift_pt_autotune": ([POINTER(Problem), POINTER(Params), c_int32, POINTER(Tuning)], c_int),
    "drift_pt_autotune_clear": ([], None),
}

_lib = None
_load_error: str | None = None


def load():
    """Load libdriftpt.so once; raises RuntimeError("... not built ...") when it is missing."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _load_error = f"GPU engine not built: {LIB_PATH} is missing (run __graft_entry__.build())"
        raise RuntimeError(_load_error)
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as exc:  # pragma: no cover - depends on the host
        _load_error = f"GPU engine not built or not loadable: {exc}"
        raise RuntimeError(_load_error) from exc
    for name, (args, ret) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ret
    if lib.drift_pt_abi_version() != ABI_VERSION:
        raise RuntimeError(
            f"GPU engine not built for this package: ABI {lib.drift_pt_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


def last_error(lib) -> str:
    msg = lib.drift_pt_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def ptr(arr, ctype):
    """Pointer to a C-contiguous numpy array (None passes NULL)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(POINTER(ctype))


def vptr(arr):
    if arr is None:
        return None
    return ctypes.c_void_p(arr.ctypes.data)
