"""Temperature ladders: geometric, feedback-optimised, and learned.

The reference runs every instance on the geometric ladder T_k = T_min (T_max/T_min)^(k/(R-1))
(drift/solvers/parallel_tempering.py:29-37).  This module adds two ways to place the same number of
rungs between the same end points; neither touches the Metropolis or exchange rules, so detailed
balance is unchanged:

* ``feedback_optimise`` -- the feedback method of Katzgraber, Trebst, Huse and Troyer
  (J. Stat. Mech. P03018, 2006).  Each replica carries a direction label (last end of the ladder it
  visited); the engine counts, per rung, the rounds spent there by replicas last seen at the cold
  end (n_up) and at the hot end (n_down).  With f = n_up / (n_up + n_down), rungs are moved so that
  every interval carries the same share of sqrt(df): in log T, the new density is
  sqrt(eta |df/du|), which makes f linear in the rung index and maximises round trips.
* ``LadderNet`` -- a 16 -> 3x128 -> 32 network (BF16 weights, FP32 accumulation, trained and run by
  the library's HIP kernels, include/drift_mlp.h) mapping 16 instance statistics to rung positions.

Ladders are carried as positions x in [0, 1] on the log-temperature axis: x_0 = 0 (T_min, coldest,
rung 0), x_{R-1} = 1 (T_max); the geometric ladder is x_j = j / (R - 1).
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int32, c_uint64, c_void_p

import numpy as np

from . import _lib

N_FEATURES = 16
N_RUNGS = 32


# ------------------------------------------------------------------------------------------------
# ladder positions <-> temperatures
# ------------------------------------------------------------------------------------------------
def geometric_positions(R: int) -> np.ndarray:
    return np.linspace(0.0, 1.0, R) if R > 1 else np.zeros(1)


def temperatures_from_positions(x, T_min: float, T_max: float) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    return np.exp(np.log(T_min) + x * (np.log(T_max) - np.log(T_min)))


def betas_from_positions(x, T_min: float, T_max: float) -> np.ndarray:
    return 1.0 / temperatures_from_positions(x, T_min, T_max)


def positions_from_temperatures(T) -> np.ndarray:
    u = np.log(np.asarray(T, dtype=np.float64))
    return (u - u[0]) / (u[-1] - u[0])


# ------------------------------------------------------------------------------------------------
# feedback optimisation
# ------------------------------------------------------------------------------------------------
def up_fraction(rung_up, rung_down) -> np.ndarray:
    """f_k = n_up / (n_up + n_down), pinned to 1 at the cold end and 0 at the hot end, made
    non-increasing (running minimum from the cold end), unvisited rungs interpolated."""
    up = np.asarray(rung_up, dtype=np.float64)
    down = np.asarray(rung_down, dtype=np.float64)
    tot = up + down
    R = up.size
    f = np.full(R, np.nan)
    seen = tot > 0
    f[seen] = up[seen] / tot[seen]
    f[0], f[-1] = 1.0, 0.0
    k = np.arange(R)
    ok = ~np.isnan(f)
    f = np.interp(k, k[ok], f[ok])
    return np.minimum.accumulate(np.clip(f, 0.0, 1.0))


def feedback_update(x, f, *, damping: float = 0.5, floor: float = 1e-4) -> np.ndarray:
    """One feedback step: interval k of the new ladder carries an equal share of sqrt(df_k)."""
    x = np.asarray(x, dtype=np.float64)
    df = np.maximum(-np.diff(np.asarray(f, dtype=np.float64)), floor)
    w = np.sqrt(df)
    W = np.concatenate([[0.0], np.cumsum(w)])
    target = np.linspace(0.0, W[-1], x.size)
    new = np.interp(target, W, x)
    new[0], new[-1] = 0.0, 1.0
    out = (1.0 - damping) * new + damping * x
    return np.maximum.accumulate(out)


def run_instances(models, *, n_replicas: int, T_min: float, T_max: float, n_rounds: int,
                  sweeps_per_round: int = 1, seeds=None, betas=None, record_history: bool = False,
                  record_trace: bool = False, targets=None):
    """PtResults of many instances in input order, in one batch call: those that fit one
    workgroup's LDS in one persistent launch, the others on the per-launch engine, four at a time
    on streams of their own.  betas: None or one ladder per instance; targets: None or one energy
    (or None) per instance to stop at."""
    from . import gpu

    m = len(models)
    seeds = list(range(m)) if seeds is None else list(seeds)
    return gpu.parallel_tempering_gpu_batch(
        list(models), n_replicas=n_replicas, T_min=T_min, T_max=T_max, n_rounds=n_rounds,
        sweeps_per_round=sweeps_per_round, seeds=seeds, betas=betas,
        record_history=record_history, record_trace=record_trace, targets=targets)


def feedback_optimise(models, *, T_min: float, T_max: float, R: int = N_RUNGS,
                      iterations: int = 4, sweeps: int = 500, seeds=None, damping: float = 0.5):
    """Feedback-optimised ladders for a list of instances, each iteration one batched launch for
    the instances that fit one workgroup (the rest one launch each).

    Iteration i runs ``sweeps * 2**i`` sweeps (one exchange attempt per sweep) on every instance
    with its current ladder, then moves the rungs.  Returns (positions (m, R), last results)."""

    m = len(models)
    x = np.tile(geometric_positions(R), (m, 1))
    seeds = list(range(m)) if seeds is None else list(seeds)
    res = None
    for it in range(iterations):
        betas = [betas_from_positions(x[i], T_min, T_max) for i in range(m)]
        res = run_instances(models, n_replicas=R, T_min=T_min, T_max=T_max,
                            n_rounds=sweeps * (1 << it), sweeps_per_round=1,
                            seeds=[s + 7919 * it for s in seeds], betas=betas)
        for i, r in enumerate(res):
            x[i] = feedback_update(x[i], up_fraction(r.rung_up, r.rung_down), damping=damping)
    return x, res


def power_positions(gamma: float, R: int = N_RUNGS) -> np.ndarray:
    """Rung positions x_j = (j / (R - 1))^gamma on the log-temperature axis: gamma = 1 is the
    geometric ladder, gamma > 1 packs rungs towards the cold end."""
    return (np.arange(R) / (R - 1)) ** float(gamma)


BUDGET_GAMMAS = (0.5, 0.7, 1.0, 1.4, 2.0)


def nearest_power_positions(x, gammas=(min(BUDGET_GAMMAS), max(BUDGET_GAMMAS))) -> np.ndarray:
    """The power ladder power_positions(gamma) closest to positions x in least squares, gamma
    restricted to [gammas[0], gammas[1]] (the span of the shapes a network was trained on), so a
    prediction outside the labelled family falls back to its nearest member."""
    x = np.asarray(x, dtype=np.float64)
    R = x.size
    grid = np.linspace(gammas[0], gammas[1], 301)
    err = [np.sum((power_positions(g, R) - x) ** 2) for g in grid]
    return power_positions(float(grid[int(np.argmin(err))]), R)


def budget_labels(models, *, T_min: float, T_max: float, R: int = N_RUNGS, budget: int = 20000,
                  gammas=BUDGET_GAMMAS, seeds=None):
    """Time-to-target labels: every instance runs each ladder shape power_positions(gamma) for
    ``budget`` sweeps on the same streams; its label is the shape with the lowest best energy,
    the one reaching that energy first on ties.  Returns (positions (m, R), chosen gamma index)."""
    m = len(models)
    seeds = list(range(m)) if seeds is None else list(seeds)
    E = np.zeros((len(gammas), m))
    first = np.zeros((len(gammas), m))
    for g, gamma in enumerate(gammas):
        b = betas_from_positions(power_positions(gamma, R), T_min, T_max)
        res = run_instances(models, n_replicas=R, T_min=T_min, T_max=T_max, n_rounds=budget,
                            sweeps_per_round=1, seeds=seeds, betas=[b] * m, record_history=True)
        for i, r in enumerate(res):
            E[g, i] = r.best_E
            first[g, i] = r.best_round
    # lexicographic: lowest energy, then earliest round
    order = np.lexsort((first, E), axis=0)[0]
    x = np.array([power_positions(gammas[k], R) for k in order])
    return x, order


def round_trip_rate(res, sweeps: int) -> float:
    """Round trips per replica per sweep of a result."""
    return float(np.sum(res.round_trips)) / (len(res.round_trips) * max(sweeps, 1))


# ------------------------------------------------------------------------------------------------
# the instance mix the network is trained and evaluated on
# ------------------------------------------------------------------------------------------------
# family: (T_min, T_max) of its ladder
MIX_FAMILIES = {"ea3d": (0.3, 3.0), "rrg": (0.1, 3.0), "er": (0.2, 10.0)}


def mix_instance(family: str, seed: int, n_target: int = 0):
    """One instance of the mix: a 3-D +/-J glass (the nearest cube), a random 3-regular graph with
    N(0,1) couplings and N(0, 0.1^2) fields, or a sparse Erdos-Renyi graph with mean degree drawn
    from [3, 12] and integer weights in {+-1, +-2, +-3}.  N is log-uniform over 256..16384, or
    about n_target when that is given."""
    from .builders.lattice import cubic_pmj_glass, er_sparse_weighted, random_regular_gaussian

    rng = np.random.default_rng(seed)
    n = n_target or int(round(np.exp(rng.uniform(np.log(256), np.log(16384)))))
    if family == "ea3d":
        return cubic_pmj_glass(max(6, int(round(n ** (1 / 3)))), seed)
    if family == "rrg":
        return random_regular_gaussian(2 * (n // 2), seed)
    if family == "er":
        return er_sparse_weighted(n, float(rng.uniform(3.0, 12.0)), seed)
    raise ValueError(f"unknown family {family!r}; expected one of {sorted(MIX_FAMILIES)}")


def heldout_instances(per_family: int, n_target: int = 65536):
    """The held-out instances: per_family of each family at about n_target spins, seeds apart from
    every training seed.  Returns [(family, model)]."""
    return [(fam, mix_instance(fam, 900_000 + 1000 * fi + k, n_target))
            for fi, fam in enumerate(MIX_FAMILIES) for k in range(per_family)]


def training_instances(per_family: int):
    """The training instances: per_family of each family, N log-uniform over 256..16384."""
    return [(fam, mix_instance(fam, 10_000 * fi + k))
            for fi, fam in enumerate(MIX_FAMILIES) for k in range(per_family)]


# ------------------------------------------------------------------------------------------------
# instance features
# ------------------------------------------------------------------------------------------------
def structural_features(model, T_min: float, T_max: float) -> np.ndarray:
    """Features 0-11: size, degree and coupling statistics, colour count and ladder span."""
    rp, ci, w = model.to_csr()
    rp = np.asarray(rp, dtype=np.int64)
    h = np.asarray(model.h, dtype=np.float64)
    deg = np.diff(rp).astype(np.float64)
    w = np.asarray(w, dtype=np.float64)
    aw = np.abs(w) if w.size else np.zeros(1)
    lib = _lib.load()
    n = model.n
    cptr = np.zeros(n + 1, dtype=np.int32)
    cs = np.zeros(n, dtype=np.int32)
    rp32 = np.ascontiguousarray(rp, dtype=np.int32)
    ci32 = np.ascontiguousarray(ci, dtype=np.int32)
    k = lib.drift_pt_greedy_coloring(n, _lib.ptr(rp32, _lib.c_int32), _lib.ptr(ci32, _lib.c_int32),
                                     _lib.ptr(cptr, _lib.c_int32), _lib.ptr(cs, _lib.c_int32))
    if k < 1:
        raise RuntimeError(f"colouring failed: {_lib.last_error(lib)}")
    return np.array([
        np.log(n), deg.mean(), deg.var(), aw.mean(), w.std() if w.size else 0.0,
        w.mean() if w.size else 0.0, aw.max() / max(aw.mean(), 1e-12),
        float((w < 0).mean()) if w.size else 0.0, np.abs(h).mean(), h.std(), float(k),
        np.log(T_max / T_min)])


def dynamic_features(res, n: int, T: np.ndarray) -> np.ndarray:
    """Features 12-15 from a short geometric-ladder run with an energy trace: specific-heat proxies
    var(E) beta^2 / n at a quarter, half and three quarters of the ladder, and the mean exchange
    acceptance."""
    tr = res.rung_energy_trace
    tail = tr[tr.shape[0] // 2:]
    R = tr.shape[1]
    out = []
    for r in (R // 4, R // 2, (3 * R) // 4):
        out.append(tail[:, r].var() / (T[r] ** 2) / n)
    acc = res.swap_accepted.sum() / max(res.swap_tried.sum(), 1)
    return np.array(out + [acc])


def instance_features(models, *, T_min: float, T_max: float, R: int = N_RUNGS,
                      short_sweeps: int = 200, seeds=None) -> np.ndarray:
    """The 16 features of each instance (one short run for the dynamic part)."""
    T = temperatures_from_positions(geometric_positions(R), T_min, T_max)
    res = run_instances(models, n_replicas=R, T_min=T_min, T_max=T_max, n_rounds=short_sweeps,
                        sweeps_per_round=1, seeds=seeds, record_trace=True)
    feats = [np.concatenate([structural_features(m, T_min, T_max), dynamic_features(r, m.n, T)])
             for m, r in zip(models, res)]
    return np.array(feats)


# ------------------------------------------------------------------------------------------------
# the network
# ------------------------------------------------------------------------------------------------
def _mlp_lib():
    lib = _lib.load()
    if not getattr(lib, "_mlp_bound", False):
        P = POINTER(c_float)
        sig = {
            "drift_mlp_create": ([c_uint64, POINTER(c_void_p)], c_int32),
            "drift_mlp_destroy": ([c_void_p], c_int32),
            "drift_mlp_get_params": ([c_void_p, P], c_int32),
            "drift_mlp_set_params": ([c_void_p, P], c_int32),
            "drift_mlp_forward": ([c_void_p, P, c_int32, P], c_int32),
            "drift_mlp_gradients": ([c_void_p, P, P, c_int32, P, P], c_int32),
            "drift_mlp_train_step": ([c_void_p, P, P, c_int32, c_float, P], c_int32),
        }
        for name, (args, ret) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ret
        lib._mlp_bound = True
    return lib


N_PARAMS = 39328


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class LadderNet:
    """Handle on the device-resident network; inputs are standardised with (mu, sd) before use."""

    def __init__(self, seed: int = 0):
        self._lib = _mlp_lib()
        h = c_void_p()
        rc = self._lib.drift_mlp_create(int(seed) & 0xFFFFFFFFFFFFFFFF, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"ladder network: {_lib.last_error(self._lib)}")
        self._h = h
        self.mu = np.zeros(N_FEATURES)
        self.sd = np.ones(N_FEATURES)

    def close(self):
        if self._h:
            self._lib.drift_mlp_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"ladder network: {_lib.last_error(self._lib)}")

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, dtype=np.float32)
        self._check(self._lib.drift_mlp_get_params(self._h, p.ctypes.data_as(POINTER(c_float))))
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        assert p.size == N_PARAMS
        self._check(self._lib.drift_mlp_set_params(self._h, p.ctypes.data_as(POINTER(c_float))))

    def fit_scaler(self, F):
        F = np.asarray(F, dtype=np.float64)
        self.mu = F.mean(axis=0)
        self.sd = np.where(F.std(axis=0) > 1e-12, F.std(axis=0), 1.0)

    def _scale(self, F):
        return _f32((np.asarray(F, dtype=np.float64) - self.mu) / self.sd)

    def raw_forward(self, X) -> np.ndarray:
        X = _f32(np.atleast_2d(X))
        out = np.zeros((X.shape[0], N_RUNGS), dtype=np.float32)
        self._check(self._lib.drift_mlp_forward(self._h, X.ctypes.data_as(POINTER(c_float)),
                                                X.shape[0], out.ctypes.data_as(POINTER(c_float))))
        return out

    def raw_gradients(self, X, T):
        X, T = _f32(np.atleast_2d(X)), _f32(np.atleast_2d(T))
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = c_float()
        self._check(self._lib.drift_mlp_gradients(
            self._h, X.ctypes.data_as(POINTER(c_float)), T.ctypes.data_as(POINTER(c_float)),
            X.shape[0], g.ctypes.data_as(POINTER(c_float)), ctypes.byref(loss)))
        return g, float(loss.value)

    def train_step(self, X, T, lr: float) -> float:
        X, T = _f32(np.atleast_2d(X)), _f32(np.atleast_2d(T))
        loss = c_float()
        self._check(self._lib.drift_mlp_train_step(
            self._h, X.ctypes.data_as(POINTER(c_float)), T.ctypes.data_as(POINTER(c_float)),
            X.shape[0], float(lr), ctypes.byref(loss)))
        return float(loss.value)

    def fit(self, F, targets, *, epochs: int = 2000, lr: float = 3e-3, batch: int = 0,
            seed: int = 0):
        """Adam on standardised features; full batch unless ``batch`` is given.  Returns losses."""
        self.fit_scaler(F)
        X = self._scale(F)
        T = _f32(targets)
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            if batch and batch < X.shape[0]:
                idx = rng.choice(X.shape[0], batch, replace=False)
                losses.append(self.train_step(X[idx], T[idx], lr))
            else:
                losses.append(self.train_step(X, T, lr))
        return np.array(losses)

    def predict(self, F) -> np.ndarray:
        """Rung positions (m, 32) for raw features F (m, 16)."""
        return self.raw_forward(self._scale(F)).astype(np.float64)

    def save(self, path: str):
        np.savez(path, params=self.params, mu=self.mu, sd=self.sd)

    def load(self, path: str):
        with np.load(path, allow_pickle=False) as d:   # plain arrays only
            self.params, self.mu, self.sd = d["params"], d["mu"], d["sd"]

    def predict_ladder(self, F) -> np.ndarray:
        """Rung positions for raw features F, each projected onto the nearest power ladder of the
        shapes the network was trained on (nearest_power_positions): a prediction for an
        instance unlike the training ones cannot leave the labelled family."""
        return np.array([nearest_power_positions(x) for x in self.predict(F)])


# ------------------------------------------------------------------------------------------------
# time to target
# ------------------------------------------------------------------------------------------------
def sweeps_to_target(best_history, target: float, tol: float = 1e-9):
    """First round (1-based count of sweeps when sweeps_per_round = 1) whose best energy reaches the
    target, or None."""
    hit = np.nonzero(np.asarray(best_history) <= target + tol * max(1.0, abs(target)))[0]
    return int(hit[0]) + 1 if hit.size else None
