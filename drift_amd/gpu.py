"""drift_amd.gpu -- the GPU parallel-tempering entry point (drop-in for drift/gpu.py).

Same surface as the reference module (drift/gpu.py:33-114): ``gpu_binary_path()``,
``gpu_available()`` and ``parallel_tempering_gpu(model, *, n_replicas=16, T_min=0.1, T_max=5.0,
n_rounds=300, sweeps_per_round=5, seed=0) -> PtResult``.  Instead of writing problem.bin and running
a CUDA executable in a subprocess, it passes the CSR arrays to libdriftpt.so (gfx950 HIP kernels)
through ctypes and reads the results back from caller-owned numpy buffers.

Couplings that are integers (or halves / quarters of integers, as QUBO conversions give) run on
the exact integer path: int8/int32 weights, integer energies, table-driven Metropolis thresholds.
Anything else runs on the float32 path.  Extra keyword arguments are additive: an explicit beta
ladder, the engine choice, the tile size and diagnostics.  There is no CPU fallback: without the
library or a gfx950 device the call raises RuntimeError("... not built ...").
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .solvers.parallel_tempering import PtResult, geometric_ladder, ladder_betas

__all__ = [
    "encode_problem",
    "gpu_available",
    "gpu_binary_path",
    "parallel_tempering_gpu",
    "parallel_tempering_gpu_batch",
]


def gpu_binary_path() -> str:
    """Absolute path of the compiled engine (libdriftpt.so)."""
    return _lib.LIB_PATH


def gpu_available() -> bool:
    """True iff the engine library loads and a gfx950 device is visible.  Never raises."""
    try:
        return bool(_lib.load().drift_pt_device_available())
    except Exception:
        return False


def _engine_lib():
    lib = _lib.load()  # raises RuntimeError("GPU engine not built ...") when missing
    if not lib.drift_pt_device_available():
        raise RuntimeError(
            f"GPU engine not built for this host: {_lib.LIB_PATH} loaded but no gfx950 device "
            "is visible")
    return lib


def _csr_of(model):
    if hasattr(model, "to_csr"):
        row_ptr, col_idx, weight = model.to_csr()
    else:
        raise TypeError(f"unsupported model type {type(model).__name__}")
    return (np.ascontiguousarray(row_ptr, dtype=np.int32),
            np.ascontiguousarray(col_idx, dtype=np.int32),
            np.asarray(weight, dtype=np.float64),
            np.asarray(model.h, dtype=np.float64))


def _integer_encoding(weight: np.ndarray, h: np.ndarray):
    """Smallest k in {0, 1, 2} such that 2^k J and 2^k h are integers below 2^30, else None."""
    w4, h4 = weight * 4.0, h * 4.0
    if (np.all(np.isfinite(w4)) and np.all(np.isfinite(h4))
            and not (np.array_equal(w4, np.round(w4)) and np.array_equal(h4, np.round(h4)))):
        return None   # 4J or 4h has a fraction, so 2^k J or 2^k h has one for every k <= 2
    for k in range(3):
        sc = float(1 << k)
        ws, hs = weight * sc, h * sc
        if not (np.all(np.isfinite(ws)) and np.all(np.isfinite(hs))):
            return None
        if np.array_equal(ws, np.round(ws)) and np.array_equal(hs, np.round(hs)):
            big = max(np.abs(ws).max(initial=0.0), np.abs(hs).max(initial=0.0))
            if big < 2.0 ** 30:
                return k, ws.astype(np.int64), hs.astype(np.int64)
    return None


class EncodedProblem:
    """numpy buffers of one instance plus the ctypes struct that points at them."""

    def __init__(self, model, coloring=None, force_float: bool = False):
        row_ptr, col_idx, weight, h = _csr_of(model)
        self.n = int(h.shape[0])
        self.row_ptr, self.col_idx = row_ptr, col_idx
        enc = None if force_float else _integer_encoding(weight, h)
        if enc is not None:
            k, wi, hi = enc
            self.scale = 1.0 / (1 << k)
            if wi.size == 0 or np.abs(wi).max() <= 127:
                self.weight_type, self.weight = _lib.W_I8, wi.astype(np.int8)
            else:
                self.weight_type, self.weight = _lib.W_I32, wi.astype(np.int32)
            self.h = hi.astype(np.int32) if np.any(hi) else None
        else:
            self.scale = 1.0
            self.weight_type = _lib.W_F32
            self.weight = weight.astype(np.float32)
            h32 = h.astype(np.float32)
            self.h = h32 if np.any(h32) else None
        if coloring is not None:
            cptr, cspins = coloring
            self.color_ptr = np.ascontiguousarray(cptr, dtype=np.int32)
            self.color_spins = np.ascontiguousarray(cspins, dtype=np.int32)
            self.n_colors = int(self.color_ptr.size - 1)
        else:
            self.color_ptr = self.color_spins = None
            self.n_colors = 0
        self.struct = _lib.Problem(
            n=self.n,
            nnz=int(col_idx.size),
            row_ptr=_lib.ptr(self.row_ptr, _lib.c_int32),
            col_idx=_lib.ptr(self.col_idx, _lib.c_int32),
            weight_type=self.weight_type,
            n_colors=self.n_colors,
            weight=_lib.vptr(self.weight),
            h=_lib.vptr(self.h),
            energy_scale=self.scale,
            color_ptr=_lib.ptr(self.color_ptr, _lib.c_int32),
            color_spins=_lib.ptr(self.color_spins, _lib.c_int32),
        )

    @property
    def is_integer(self) -> bool:
        return self.weight_type != _lib.W_F32


def encode_problem(model, coloring=None, force_float: bool = False) -> EncodedProblem:
    return EncodedProblem(model, coloring, force_float)


class _Buffers:
    """Caller-owned result arrays for one run."""

    def __init__(self, n: int, R: int, n_rounds: int, final: bool, history: bool,
                 trace: bool = False):
        self.best_s = np.zeros(n, dtype=np.int8)
        self.tried = np.zeros(max(R - 1, 1), dtype=np.int64)
        self.acc = np.zeros(max(R - 1, 1), dtype=np.int64)
        self.trips = np.zeros(R, dtype=np.int32)
        self.up = np.zeros(R, dtype=np.int64)
        self.down = np.zeros(R, dtype=np.int64)
        self.rung_of = np.zeros(R, dtype=np.int32)
        self.final_E = np.zeros(R, dtype=np.int64)
        self.final_spins = np.zeros((R, n), dtype=np.int8) if final else None
        self.history = np.zeros(max(n_rounds, 1), dtype=np.float64) if history else None
        self.trace = np.zeros((max(n_rounds, 1), R), dtype=np.float64) if trace else None
        self.struct = _lib.Result(
            best_s=_lib.ptr(self.best_s, _lib.c_int8),
            swap_tried=_lib.ptr(self.tried, _lib.c_int64),
            swap_accepted=_lib.ptr(self.acc, _lib.c_int64),
            round_trips=_lib.ptr(self.trips, _lib.c_int32),
            rung_up=_lib.ptr(self.up, _lib.c_int64),
            rung_down=_lib.ptr(self.down, _lib.c_int64),
            rung_of=_lib.ptr(self.rung_of, _lib.c_int32),
            final_E_units=_lib.ptr(self.final_E, _lib.c_int64),
            final_spins=_lib.ptr(self.final_spins, _lib.c_int8),
            best_history=_lib.ptr(self.history, _lib.c_double),
            rung_energy_trace=_lib.ptr(self.trace, _lib.c_double),
        )


def _params(n_replicas, T_min, T_max, n_rounds, sweeps_per_round, seed, betas, engine,
            block_threads, final, history, trace=False, time_kernels=False, graph=True,
            autotune=False, plain_launch=False, target=None, slots_per_group=0,
            couplings_in_lds=False, parts_per_group=0):
    flags = ((_lib.RECORD_FINAL if final else 0) | (_lib.RECORD_HISTORY if history else 0)
             | (_lib.RECORD_TRACE if trace else 0) | (_lib.TIME_KERNELS if time_kernels else 0)
             | (0 if graph else _lib.NO_GRAPH) | (_lib.AUTOTUNE if autotune else 0)
             | (_lib.PLAIN_LAUNCH if plain_launch else 0)
             | (_lib.STOP_AT_TARGET if target is not None else 0)
             | (_lib.COUPLINGS_LDS if couplings_in_lds else 0))
    if engine not in _lib.ENGINES:
        raise ValueError(f"engine must be one of {sorted(_lib.ENGINES)}")
    return _lib.Params(
        seed=int(seed) & 0xFFFFFFFFFFFFFFFF,
        n_replicas=int(n_replicas),
        n_rounds=int(n_rounds),
        sweeps_per_round=int(sweeps_per_round),
        engine=_lib.ENGINES[engine],
        T_min=float(T_min),
        T_max=float(T_max),
        beta=_lib.ptr(betas, _lib.c_double),
        flags=flags,
        block_threads=int(block_threads),
        target_energy=0.0 if target is None else float(target),
        slots_per_group=int(slots_per_group),
        parts_per_group=int(parts_per_group),
    )


def _ladder(n_replicas, T_min, T_max, betas):
    if betas is None:
        temps = geometric_ladder(n_replicas, T_min, T_max)
        b = ladder_betas(temps)
    else:
        b = np.ascontiguousarray(betas, dtype=np.float64)
        if b.shape != (n_replicas,):
            raise ValueError(f"betas must have n_replicas={n_replicas} entries")
        temps = 1.0 / b
    return temps, np.ascontiguousarray(b, dtype=np.float64)


def _to_result(res: _lib.Result, buf: _Buffers, temps, n_rounds, R, history) -> PtResult:
    tried = int(buf.tried[: R - 1].sum()) if R > 1 else 0
    accepted = int(buf.acc[: R - 1].sum()) if R > 1 else 0
    secs = float(res.seconds)
    return PtResult(
        best_s=buf.best_s.astype(np.float64),
        best_E=float(res.best_E),
        temperatures=temps,
        swap_rate=accepted / tried if tried else 0.0,
        best_history=buf.history[:n_rounds].copy() if history else None,
        throughput=float(res.attempts) / secs if secs > 0 else None,
        seconds=secs,
        swap_acceptance=(buf.acc[: R - 1] / np.maximum(buf.tried[: R - 1], 1)) if R > 1 else np.zeros(0),
        round_trips=buf.trips.copy(),
        rung_up=buf.up.copy(),
        rung_down=buf.down.copy(),
        rung_of=buf.rung_of.copy(),
        final_E=buf.final_E.copy(),
        final_spins=buf.final_spins,
        best_round=int(res.best_round),
        best_rung=int(res.best_rung),
        best_E_units=int(res.best_E_units),
        engine=_lib.ENGINE_NAMES.get(int(res.engine_used), str(res.engine_used)),
        rung_energy_trace=None if buf.trace is None else buf.trace[:n_rounds].copy(),
        swap_tried=buf.tried[: max(R - 1, 0)].copy(),
        swap_accepted=buf.acc[: max(R - 1, 0)].copy(),
        kernel_seconds=float(res.kernel_seconds),
        kernel_attempts=float(res.kernel_attempts),
        kernel_launches=int(res.kernel_launches),
        other_kernel_seconds=float(res.other_kernel_seconds),
        other_kernel_launches=int(res.other_kernel_launches),
        rounds_run=int(res.rounds_run),
        block_threads=int(res.block_threads_used),
        slots_per_group=int(res.slots_per_group_used),
        couplings_in_lds=bool(res.couplings_in_lds_used),
        parts_per_group=int(res.parts_per_group_used),
    )


def parallel_tempering_gpu(
    model,
    *,
    n_replicas: int = 16,
    T_min: float = 0.1,
    T_max: float = 5.0,
    n_rounds: int = 300,
    sweeps_per_round: int = 5,
    seed: int = 0,
    betas=None,
    engine: str = "auto",
    block_threads: int = 0,
    coloring=None,
    record_history: bool = True,
    return_final: bool = False,
    record_trace: bool = False,
    time_kernels: bool = False,
    graph: bool = True,
    force_float: bool = False,
    autotune: bool = False,
    plain_launch: bool = False,
    target_energy: float | None = None,
    slots_per_group: int = 0,
    couplings_in_lds: bool = False,
    parts_per_group: int = 0,
) -> PtResult:
    """Solve ``model`` by parallel tempering on the GPU; same signature and result as the CPU
    ``parallel_tempering`` (drift/gpu.py:43-114).  ``throughput`` is spin-flip attempts per second
    and ``seconds`` the device time of the sweep and exchange loop.  ``autotune`` (with
    ``engine="auto"``) takes the engine and block size from the library's autotune cache, timing
    the candidates the first time this shape class is seen; results do not depend on it.
    ``plain_launch`` starts the persistent engines with an ordinary dispatch (for profilers that
    cannot follow cooperative launches; needs the device to itself).  ``target_energy`` stops the
    run after the first round end whose best energy is at or below it (engines lds, global and
    lattice); ``rounds_run`` then says how many rounds ran, and the result is that of a run of
    that many rounds.  ``slots_per_group`` sets the split engine's replica slots per workgroup
    (a power of two up to 64; 0: the library's choice); ``couplings_in_lds`` has it stage the
    couplings in LDS too (int8 couplings with |w| <= 7, n <= 4096); ``parts_per_group`` splits
    each group's positions over that many workgroups (1 to 64; 0: the library's choice)."""
    lib = _engine_lib()
    temps, b = _ladder(n_replicas, T_min, T_max, betas)
    prob = encode_problem(model, coloring, force_float)
    prm = _params(n_replicas, T_min, T_max, n_rounds, sweeps_per_round, seed, b, engine,
                  block_threads, return_final, record_history, record_trace, time_kernels,
                  graph, autotune, plain_launch, target_energy, slots_per_group,
                  couplings_in_lds, parts_per_group)
    buf = _Buffers(prob.n, n_replicas, n_rounds, return_final, record_history, record_trace)
    res = buf.struct
    rc = lib.drift_pt_run(prob.struct, prm, res)
    if rc != 0:
        raise RuntimeError(f"GPU engine failed (status {rc}): {_lib.last_error(lib)}")
    return _to_result(res, buf, temps, n_rounds, n_replicas, record_history)


def parallel_tempering_gpu_batch(
    models,
    *,
    n_replicas: int = 16,
    T_min: float = 0.1,
    T_max: float = 5.0,
    n_rounds: int = 300,
    sweeps_per_round: int = 5,
    seeds=None,
    betas=None,
    record_history: bool = False,
    return_final: bool = False,
    record_trace: bool = False,
    force_float: bool = False,
    block_threads: int = 0,
    no_alloc: bool = False,
    targets=None,
    slots_per_group: int = 0,
    autotune: bool = False,
    plain_launch: bool = False,
    couplings_in_lds: bool = False,
    parts_per_group=0,
) -> list[PtResult]:
    """Many independent instances in one persistent launch.  ``seeds`` (default 0, 1, 2, ...) and
    ``betas`` (one ladder per instance, or None for the geometric ladder) are per instance; each
    result equals ``parallel_tempering_gpu`` on that instance alone.  ``block_threads`` sets the
    workgroup size (0: the library's choice).  ``no_alloc`` runs the batch through
    drift_pt_run_batch_noalloc, which allocates no device memory (a fixed region of
    ``_lib.NOALLOC_BYTES`` per device; larger batches raise RuntimeError); the results are the
    same.  Instances too large for one workgroup run on the split engine in a second persistent
    launch when their replica groups fill the chip (``slots_per_group`` sets the group size, 0:
    the library's choice), else on the per-launch engine, four at a time on streams of their own.
    ``targets`` (one energy or None per instance) stops each instance at its target, as
    ``target_energy`` does for one.  ``autotune`` times the launch shapes of each persistent
    launch on the batch itself (block size; for the split engine also slots per group), keeps the
    choice for the batch's shape class, and runs with it; results do not depend on it.
    ``plain_launch`` starts the split engine's persistent launch with an ordinary dispatch (for
    profilers that cannot follow cooperative launches; needs the device to itself);
    ``couplings_in_lds`` has the split launch stage every instance's couplings in LDS;
    ``parts_per_group`` (one count, or one per instance) splits replica groups' positions over that
    many workgroups."""
    lib = _engine_lib()
    models = list(models)
    m = len(models)
    if m == 0:
        return []
    seeds = list(range(m)) if seeds is None else [int(s) for s in seeds]
    parts = ([int(parts_per_group)] * m if np.isscalar(parts_per_group)
             else [int(p) for p in parts_per_group])
    if len(parts) != m:
        raise ValueError(f"parts_per_group must be one count or {m} counts")
    probs, prms, bufs, ladders = [], [], [], []
    for i, model in enumerate(models):
        temps, b = _ladder(n_replicas, T_min, T_max, None if betas is None else betas[i])
        ladders.append((temps, b))
        probs.append(encode_problem(model, None, force_float))
        prms.append(_params(n_replicas, T_min, T_max, n_rounds, sweeps_per_round, seeds[i], b,
                            "lds", block_threads, return_final, record_history, record_trace,
                            target=None if targets is None else targets[i],
                            slots_per_group=slots_per_group, autotune=autotune,
                            plain_launch=plain_launch, couplings_in_lds=couplings_in_lds,
                            parts_per_group=parts[i]))
        bufs.append(_Buffers(probs[-1].n, n_replicas, n_rounds, return_final, record_history,
                             record_trace))
    P = (_lib.Problem * m)(*[p.struct for p in probs])
    Q = (_lib.Params * m)(*prms)
    Rs = (_lib.Result * m)(*[bf.struct for bf in bufs])
    rc = (lib.drift_pt_run_batch_noalloc if no_alloc else lib.drift_pt_run_batch)(P, Q, m, Rs)
    if rc != 0:
        raise RuntimeError(f"GPU engine failed (status {rc}): {_lib.last_error(lib)}")
    return [_to_result(Rs[i], bufs[i], ladders[i][0], n_rounds, n_replicas, record_history)
            for i in range(m)]


def autotune(model, *, n_replicas: int = 16, T_min: float = 0.1, T_max: float = 5.0,
             sweeps_per_round: int = 5, seed: int = 0, betas=None, coloring=None,
             trial_rounds: int = 0, force_float: bool = False) -> dict:
    """Time every engine and launch shape that can run ``model`` (drift_pt_autotune), cache the
    choice for its shape class, and return {"engine", "block_threads", "slots_per_group",
    "couplings_in_lds", "parts_per_group", "trial_rounds", "candidates": [(engine, block,
    slots_per_group, couplings_in_lds, parts_per_group, seconds), ...]} (block 0: the engine's
    own choice; seconds < 0: the candidate could not run)."""
    lib = _engine_lib()
    temps, b = _ladder(n_replicas, T_min, T_max, betas)
    prob = encode_problem(model, coloring, force_float)
    prm = _params(n_replicas, T_min, T_max, max(int(trial_rounds), 1), sweeps_per_round, seed, b,
                  "auto", 0, False, False)
    tab = _lib.Tuning()
    rc = lib.drift_pt_autotune(prob.struct, prm, int(trial_rounds), tab)
    if rc != 0:
        raise RuntimeError(f"GPU engine failed (status {rc}): {_lib.last_error(lib)}")
    cands = [(_lib.ENGINE_NAMES.get(int(tab.cand_engine[i]), str(tab.cand_engine[i])),
              int(tab.cand_block[i]), int(tab.cand_group[i]), bool(tab.cand_lds_couplings[i]),
              int(tab.cand_parts[i]), float(tab.cand_seconds[i]))
             for i in range(tab.n_candidates)]
    return {"engine": _lib.ENGINE_NAMES.get(int(tab.engine), str(tab.engine)),
            "block_threads": int(tab.block_threads), "slots_per_group": int(tab.slots_per_group),
            "couplings_in_lds": bool(tab.couplings_in_lds),
            "parts_per_group": int(tab.parts_per_group), "trial_rounds": int(tab.trial_rounds),
            "candidates": cands}


def autotune_clear() -> None:
    """Empty the library's autotune cache."""
    _engine_lib().drift_pt_autotune_clear()
