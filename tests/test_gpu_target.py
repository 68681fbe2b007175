"""Stopping at a target energy, and batches of instances too large for one workgroup.

DRIFT_PT_STOP_AT_TARGET ends a run after the first round end whose best energy is at or below the
target; the result must then be, bit for bit, the oracle's run of exactly that many rounds (best
state and round, final spins and energies, rung permutation, swap and round-trip counters).  The
target is taken from a full run's own best-energy history, so the stop round is known in advance.
Large instances in a batch run on the per-launch engine, four at a time on streams of their own;
each must equal its run alone and the oracle, with and without targets.
"""

import os

import numpy as np
import pytest

import oracle
from drift_amd import gpu
from drift_amd.builders.lattice import (cubic_pmj_glass, er_sparse_weighted,
                                        random_regular_gaussian)
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

THREADS = min(16, len(os.sched_getaffinity(0)))


def _announce(name):
    print(f"[gpu-target] {name}", flush=True)


def _oracle(model, *, seed, R, n_rounds, S, beta):
    enc = gpu.encode_problem(model)
    k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    return oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=n_rounds,
                                 sweeps_per_round=S, beta=beta, n_threads=THREADS)


def _equal(g, o):
    assert g.best_E_units == o["best_units"]
    assert (g.best_round, g.best_rung) == (o["best_round"], o["best_rung"])
    assert np.array_equal(g.best_s.astype(np.int8), o["best_s"])
    assert np.array_equal(g.final_spins, o["final_spins"])
    assert np.array_equal(g.final_E, o["final_units"])
    assert np.array_equal(g.rung_of, o["rung_of"])
    assert np.array_equal(g.swap_tried, o["swap_tried"])
    assert np.array_equal(g.swap_accepted, o["swap_acc"])
    assert np.array_equal(g.round_trips, o["round_trips"])
    assert np.array_equal(g.rung_up, o["rung_up"]) and np.array_equal(g.rung_down, o["rung_down"])


CASES = {
    "lds": (lambda: random_regular_gaussian(300, 5), 16, 1),
    "global-words": (lambda: er_sparse_weighted(6000, 6.0, 3), 32, 1),
    "global-bytes": (lambda: er_sparse_weighted(4000, 5.0, 4), 48, 2),
    "lattice": (lambda: cubic_pmj_glass(12, 6), 16, 1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_stop_at_target_equals_a_run_of_that_many_rounds(case):
    make, R, S = CASES[case]
    engine = case.split("-")[0]
    model = make()
    beta = ladder_betas(geometric_ladder(R, 0.2, 4.0))
    kw = dict(n_replicas=R, betas=beta, sweeps_per_round=S, seed=11, engine=engine,
              return_final=True, record_history=True)
    full = gpu.parallel_tempering_gpu(model, n_rounds=400, **kw)
    # a target first reached somewhere in the middle of the run
    hist = full.best_history
    target = float(hist[len(hist) // 3])
    stop = int(np.nonzero(hist <= target)[0][0]) + 1
    _announce(f"{case}: target {target} first reached after round {stop} of 400")
    g = gpu.parallel_tempering_gpu(model, n_rounds=400, target_energy=target, **kw)
    assert g.engine == engine and g.rounds_run == stop
    assert g.best_E <= target
    _equal(g, _oracle(model, seed=11, R=R, n_rounds=stop, S=S, beta=beta))
    assert np.array_equal(g.best_history[:stop], hist[:stop])
    assert np.all(g.best_history[stop:] == g.best_E)
    # a target no state reaches runs every round; one the initial states reach runs none
    g = gpu.parallel_tempering_gpu(model, n_rounds=60, target_energy=-1e9, **kw)
    assert g.rounds_run == 60
    _equal(g, _oracle(model, seed=11, R=R, n_rounds=60, S=S, beta=beta))
    g = gpu.parallel_tempering_gpu(model, n_rounds=60, target_energy=1e9, **kw)
    assert g.rounds_run == 0 and g.best_round == 0


def test_autotune_with_a_target_keeps_to_engines_that_stop():
    """A tuned run with a target: the tuner times only the engines that can stop at a target
    (lds, split, global, lattice), under a shape class apart from the untargeted one, and the
    result is the oracle's run of the rounds it took."""
    gpu.autotune_clear()
    model = cubic_pmj_glass(16, 2)
    R, S = 16, 1
    beta = ladder_betas(geometric_ladder(R, 0.2, 4.0))
    kw = dict(n_replicas=R, betas=beta, sweeps_per_round=S, seed=13, return_final=True,
              record_history=True, autotune=True)
    free = gpu.parallel_tempering_gpu(model, n_rounds=300, **kw)
    hist = free.best_history
    target = float(hist[len(hist) // 3])
    stop = int(np.nonzero(hist <= target)[0][0]) + 1
    _announce(f"tuned with a target: untargeted choice {free.engine}, stop after round {stop}")
    g = gpu.parallel_tempering_gpu(model, n_rounds=300, target_energy=target, **kw)
    assert g.engine in ("lds", "split", "global", "lattice") and g.rounds_run == stop
    _equal(g, _oracle(model, seed=13, R=R, n_rounds=stop, S=S, beta=beta))
    # the untargeted choice is still the cached one
    assert gpu.parallel_tempering_gpu(model, n_rounds=20, **kw).engine == free.engine


def test_persistent_engines_refuse_a_target():
    _announce("pslab and coop refuse a target")
    for engine, model in (("pslab", cubic_pmj_glass(32, 1)), ("coop", er_sparse_weighted(3000, 6.0, 1))):
        with pytest.raises(RuntimeError, match="STOP_AT_TARGET"):
            gpu.parallel_tempering_gpu(model, n_replicas=16, n_rounds=10, engine=engine,
                                       target_energy=0.0)


def test_batch_with_large_instances_on_streams():
    """Three instances that fit one workgroup and six that do not (the per-launch engine, four at
    a time on streams of their own: integer and float CSR instances and a 20^3 lattice), with and
    without targets: each equals its run alone and the oracle."""
    R, S, rounds = 32, 1, 120
    models = ([random_regular_gaussian(200 + 40 * i, 100 + i) for i in range(3)]
              + [er_sparse_weighted(6000 + 700 * i, 6.0, 200 + i) for i in range(4)]
              + [cubic_pmj_glass(20, 7), random_regular_gaussian(6000, 9)])
    seeds = [31 + i for i in range(len(models))]
    _announce(f"batch of {len(models)} instances, N {min(m.n for m in models)}.."
              f"{max(m.n for m in models)}, R={R}")
    kw = dict(n_replicas=R, T_min=0.25, T_max=5.0, n_rounds=rounds, sweeps_per_round=S,
              seeds=seeds, return_final=True, record_history=True)
    res = gpu.parallel_tempering_gpu_batch(models, **kw)
    beta = ladder_betas(geometric_ladder(R, 0.25, 5.0))
    for m, seed, r in zip(models, seeds, res):
        _equal(r, _oracle(m, seed=seed, R=R, n_rounds=rounds, S=S, beta=beta))
    alone = gpu.parallel_tempering_gpu(models[5], n_replicas=R, T_min=0.25, T_max=5.0,
                                       n_rounds=rounds, sweeps_per_round=S, seed=seeds[5],
                                       return_final=True)
    assert alone.engine == res[5].engine
    assert np.array_equal(alone.final_spins, res[5].final_spins)
    # every instance stops at the best energy its full run had after round 40
    targets = [float(r.best_history[39]) for r in res]
    stopped = gpu.parallel_tempering_gpu_batch(models, targets=targets, **kw)
    for m, seed, r, full, t in zip(models, seeds, stopped, res, targets):
        want = int(np.nonzero(full.best_history <= t)[0][0]) + 1
        assert r.rounds_run == want
        _equal(r, _oracle(m, seed=seed, R=R, n_rounds=want, S=S, beta=beta))
