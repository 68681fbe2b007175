"""GPU parity of the batched persistent launch (drift_pt_run_batch) at the shape of config C4.

A batch of 64 or more instances takes its own launch shape: 256-thread workgroups that sweep a
colour class in several passes, and instances prepared on several host threads.  These tests put
that path under the oracle:

* uneven 3-regular float instances with N over the whole range [64, 2048] (several at or near
  2048), R = 16, compared instance by instance with the CPU oracle bit for bit: best energy
  units, best round and rung, best state, final spins and energies, rung permutation, swap
  counters, round trips, rung occupation counters and the best-energy history;
* the full C4 batch (4096 instances) checked through properties the oracle cannot reach at that
  size: every slot's tracked energy against an fp64 recomputation from its final spins (within
  1e-5 relative), the reported best energy against the energy of the reported state, and a few
  instances run alone against the same instances inside the batch.

The reference has one instance per process (drift/gpu.py:97, cuda/ising_pt.cu:214-218); the batch
replaces that loop.
"""

import os

import numpy as np
import pytest

import oracle
from drift_amd import gpu
from drift_amd.builders.lattice import random_regular_gaussian
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

THREADS = min(16, len(os.sched_getaffinity(0)))
FIX36 = 2.0 ** -36


def _announce(name):
    print(f"[gpu-batch] {name}", flush=True)


def _uneven_sizes(count, seed, near_top=()):
    """N uniform over the even values of [64, 2048], plus the given sizes."""
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in 2 * rng.integers(32, 1025, size=count - len(near_top))]
    return sizes + [int(x) for x in near_top]


def _oracle_equal(model, r, *, seed, R, n_rounds, S, beta):
    enc = gpu.encode_problem(model)
    k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=n_rounds,
                              sweeps_per_round=S, beta=beta, n_threads=THREADS)
    assert r.best_E_units == o["best_units"]
    assert (r.best_round, r.best_rung) == (o["best_round"], o["best_rung"])
    assert np.array_equal(r.best_s.astype(np.int8), o["best_s"])
    assert np.array_equal(r.final_spins, o["final_spins"])
    assert np.array_equal(r.final_E, o["final_units"])
    assert np.array_equal(r.rung_of, o["rung_of"])
    assert np.array_equal(r.swap_tried, o["swap_tried"])
    assert np.array_equal(r.swap_accepted, o["swap_acc"])
    assert np.array_equal(r.round_trips, o["round_trips"])
    assert np.array_equal(r.rung_up, o["rung_up"]) and np.array_equal(r.rung_down, o["rung_down"])
    assert np.array_equal(r.best_history, o["best_history"])


@pytest.mark.parametrize("case", [(96, 16, 30, 1), (72, 12, 20, 3), (40, 16, 30, 1)])
def test_uneven_batch_up_to_2048_equals_oracle(case):
    """(instances, R, rounds, sweeps per round): 96 and 72 instances take the launch shape of
    C4 (64 or more instances); 40 instances with N up to 2048 take the small-batch shape
    (workgroups sized to the largest colour class)."""
    count, R, n_rounds, S = case
    sizes = _uneven_sizes(count, seed=count, near_top=(2048, 2046, 2042, 2030, 2000))
    _announce(f"batch of {count} uneven float instances, N {min(sizes)}..{max(sizes)}, R={R}, "
              f"{n_rounds} rounds x {S} sweeps")
    models = [random_regular_gaussian(n, seed=20_000 + 7 * count + i) for i, n in enumerate(sizes)]
    seeds = [5 + 11 * i for i in range(count)]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True)
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    for model, seed, r in zip(models, seeds, res):
        assert r.engine == "lds"
        _oracle_equal(model, r, seed=seed, R=R, n_rounds=n_rounds, S=S, beta=beta)


def test_degenerate_instances_in_a_batch_equal_oracle():
    """One spin, spins with no couplings at all (nnz = 0; integer and float fields), an all-zero
    model and one coupled pair beside isolated spins, batched with two ordinary instances (one
    launch per weight type), each equal to the oracle and at its closed-form ground energy."""
    from drift_amd.ising import IsingModel

    _announce("degenerate instances in a batch")
    rng = np.random.default_rng(8)
    pair = np.zeros((5, 5))
    pair[1, 3] = pair[3, 1] = -2.0
    odd = [IsingModel(np.zeros((1, 1)), np.array([2.0])),
           IsingModel(np.zeros((7, 7)), rng.normal(size=7)),
           IsingModel(np.zeros((9, 9)), rng.integers(-3, 4, size=9).astype(float)),
           IsingModel(np.zeros((5, 5)), np.zeros(5)),
           IsingModel(pair, np.zeros(5))]
    models = odd + [random_regular_gaussian(64, seed=3), random_regular_gaussian(130, seed=4)]
    R, n_rounds, S = 8, 15, 2
    seeds = [40 + i for i in range(len(models))]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True)
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    for model, seed, r in zip(models, seeds, res):
        _oracle_equal(model, r, seed=seed, R=R, n_rounds=n_rounds, S=S, beta=beta)
    for model, r in zip(odd, res):
        assert r.best_E == pytest.approx(-np.abs(model.h).sum() - np.abs(model.J).sum() / 2,
                                         rel=1e-6, abs=1e-9)


def test_maximum_replica_count_batch_equals_oracle():
    """R = 1024 in a batch: two instances too large for one workgroup's LDS at that R go to the
    split engine (each instance's groups no more than the compute units, so they are resident
    together), one small instance to the LDS engine; each equals the oracle."""
    _announce("batch at R = 1024")
    models = [random_regular_gaussian(300, seed=31), random_regular_gaussian(250, seed=32),
              random_regular_gaussian(64, seed=33)]
    R, n_rounds, S = 1024, 6, 1
    seeds = [90, 91, 92]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True)
    assert [r.engine for r in res] == ["split", "split", "lds"]
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    for model, seed, r in zip(models, seeds, res):
        _oracle_equal(model, r, seed=seed, R=R, n_rounds=n_rounds, S=S, beta=beta)


def test_full_c4_batch_properties():
    """The C4 batch itself: 4096 random 3-regular instances, N uniform over even values in
    [64, 2048], J ~ N(0,1) and h ~ N(0, 0.1^2) in float32, R = 16, beta 0.2 .. 5.0, one
    persistent launch of 300 sweeps.  The fixed-point energies the kernels track stay within
    1e-5 relative of an fp64 recomputation for every slot of every instance; each reported best
    energy is the energy of its reported state; instances run alone equal their batch results."""
    count, R, n_rounds = 4096, 16, 300
    rng = np.random.default_rng(0)
    sizes = 2 * rng.integers(32, 1025, size=count)
    _announce(f"full C4 batch: {count} instances, {int(sizes.sum())} spins, {n_rounds} sweeps")
    models = [random_regular_gaussian(int(n), seed=10_000 + i) for i, n in enumerate(sizes)]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=1,
                                           return_final=True)
    worst = 0.0
    for model, r in zip(models, res):
        e64 = model.energy_batch(r.final_spins.astype(np.float64))
        tracked = r.final_E.astype(np.float64) * FIX36
        scale = np.maximum(np.abs(e64), 1.0)
        worst = max(worst, float(np.max(np.abs(tracked - e64) / scale)))
        eb = model.energy(r.best_s)
        assert abs(r.best_E - eb) <= 1e-5 * max(abs(eb), 1.0)
        assert r.best_E_units <= int(r.final_E.min())   # the last round end is scanned too
    _announce(f"largest relative energy deviation {worst:.3e}")
    assert worst <= 1e-5
    # the largest instance, the smallest and one in the middle, alone
    order = np.argsort(sizes)
    for i in (int(order[-1]), int(order[0]), int(order[count // 2])):
        alone = gpu.parallel_tempering_gpu(models[i], n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=1, seed=i,
                                           engine="lds", return_final=True)
        assert alone.best_E_units == res[i].best_E_units
        assert np.array_equal(alone.best_s, res[i].best_s)
        assert np.array_equal(alone.final_spins, res[i].final_spins)
        assert np.array_equal(alone.rung_of, res[i].rung_of)


def test_batch_autotune_keeps_results():
    """The batch autotuner times the LDS launch's block sizes on the batch itself and runs with
    its choice; the results equal the untuned batch's bit for bit, and the oracle's."""
    count, R, n_rounds = 80, 16, 12
    sizes = _uneven_sizes(count, seed=5, near_top=(2048, 2040))
    _announce(f"batch autotune, {count} instances")
    gpu.autotune_clear()
    models = [random_regular_gaussian(n, seed=30_000 + i) for i, n in enumerate(sizes)]
    kw = dict(n_replicas=R, T_min=0.2, T_max=5.0, n_rounds=n_rounds, sweeps_per_round=1,
              return_final=True, record_history=True)
    plain = gpu.parallel_tempering_gpu_batch(models, **kw)
    tuned = gpu.parallel_tempering_gpu_batch(models, autotune=True, **kw)
    assert tuned[0].block_threads in (256, 512, 1024)
    for a, b in zip(plain, tuned):
        assert a.best_E_units == b.best_E_units and np.array_equal(a.best_s, b.best_s)
        assert np.array_equal(a.final_spins, b.final_spins) and np.array_equal(a.rung_of, b.rung_of)
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    for i in (0, count - 1):
        _oracle_equal(models[i], tuned[i], seed=i, R=R, n_rounds=n_rounds, S=1, beta=beta)


def test_batch_autotune_split_launch_keeps_results():
    """Large instances go to the split engine; the tuner picks a group size and block size on
    the batch and the results equal those of the library's default launch shape."""
    count, R, n_rounds = 12, 32, 6
    _announce(f"batch autotune on the split engine, {count} instances")
    gpu.autotune_clear()
    models = [random_regular_gaussian(12000 + 64 * i, seed=40_000 + i) for i in range(count)]
    kw = dict(n_replicas=R, T_min=0.2, T_max=5.0, n_rounds=n_rounds, sweeps_per_round=2,
              return_final=True)
    plain = gpu.parallel_tempering_gpu_batch(models, **kw)
    tuned = gpu.parallel_tempering_gpu_batch(models, autotune=True, **kw)
    assert plain[0].engine == tuned[0].engine == "split"
    for a, b in zip(plain, tuned):
        assert a.best_E_units == b.best_E_units and np.array_equal(a.best_s, b.best_s)
        assert np.array_equal(a.final_spins, b.final_spins) and np.array_equal(a.rung_of, b.rung_of)
