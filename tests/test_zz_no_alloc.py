"""The allocation-free batch entry point (drift_pt_run_batch_noalloc) against drift_pt_run_batch.

Both take the same host arrays and return the same results bit for bit; the allocation-free one
carves every device buffer from a fixed per-device region instead of allocating.  Two calls in a
row share that region, the second with a smaller batch of another weight type, so anything the
first call left behind (work queue, statistics blocks, best states) would show up in the second.
A batch larger than the region is refused with a clear error before anything is launched, and the
next call still runs.
"""

import numpy as np
import pytest

from drift_amd import _lib, gpu
from drift_amd.builders.lattice import er_sparse_weighted, random_regular_gaussian

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]


def _announce(name):
    print(f"[no-alloc] {name}", flush=True)


def _same(a, b):
    assert a.best_E_units == b.best_E_units
    assert (a.best_round, a.best_rung) == (b.best_round, b.best_rung)
    assert np.array_equal(a.best_s, b.best_s)
    assert np.array_equal(a.final_spins, b.final_spins)
    assert np.array_equal(a.final_E, b.final_E)
    assert np.array_equal(a.rung_of, b.rung_of)
    assert np.array_equal(a.swap_tried, b.swap_tried)
    assert np.array_equal(a.swap_accepted, b.swap_accepted)
    assert np.array_equal(a.round_trips, b.round_trips)
    assert np.array_equal(a.rung_up, b.rung_up) and np.array_equal(a.rung_down, b.rung_down)
    assert np.array_equal(a.best_history, b.best_history)


def _both(models, **kw):
    ref = gpu.parallel_tempering_gpu_batch(models, return_final=True, record_history=True, **kw)
    got = gpu.parallel_tempering_gpu_batch(models, return_final=True, record_history=True,
                                           no_alloc=True, **kw)
    for a, b in zip(got, ref):
        _same(a, b)
    return got


def test_no_alloc_batch_equals_the_allocating_batch_twice_in_a_row():
    rng = np.random.default_rng(3)
    sizes = [int(x) for x in 2 * rng.integers(32, 1025, size=78)] + [2048, 2046]
    _announce(f"call 1: {len(sizes)} float instances, N {min(sizes)}..{max(sizes)}, R=16")
    first = [random_regular_gaussian(n, seed=500 + i) for i, n in enumerate(sizes)]
    _both(first, n_replicas=16, T_min=0.2, T_max=5.0, n_rounds=30, sweeps_per_round=1,
          seeds=[7 * i + 1 for i in range(len(first))])
    _announce("call 2: 12 smaller int8 instances, R=8")
    second = [er_sparse_weighted(int(n), 5.0, 900 + i)
              for i, n in enumerate(rng.integers(40, 400, size=12))]
    _both(second, n_replicas=8, T_min=0.3, T_max=4.0, n_rounds=45, sweeps_per_round=2,
          seeds=[3 * i + 2 for i in range(len(second))])


def test_no_alloc_batch_refuses_more_than_its_region():
    _announce("a batch larger than the region is refused")
    models = [random_regular_gaussian(64, seed=i) for i in range(10)]
    rounds = _lib.NOALLOC_BYTES // (8 * 8)   # ten best-energy histories of 8 bytes per round
    with pytest.raises(RuntimeError, match="more than the"):
        gpu.parallel_tempering_gpu_batch(models, n_replicas=4, n_rounds=rounds, sweeps_per_round=0,
                                         record_history=True, no_alloc=True)
    _announce("the next call runs")
    _both(models[:3], n_replicas=4, T_min=0.3, T_max=3.0, n_rounds=20, sweeps_per_round=1)
