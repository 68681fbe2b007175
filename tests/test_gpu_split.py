"""GPU parity of the split engine (DRIFT_PT_ENGINE_SPLIT).

The split engine holds an instance's replica slots in groups of ``slots_per_group``, each group's
spins in one workgroup's LDS; after every round's sweeps the groups trade slot energies through
round-tagged mailboxes and all take the same round-end decisions.  It runs the one chain every
engine runs, so each case is compared with the CPU oracle bit for bit: best energy units, best
round and rung, best state, final spins and energies, rung permutation, swap counters, round
trips, rung occupation counters, best-energy history and the per-rung energy trace.

Cases cover every group size from one slot to the whole ladder, R above 64 (the workgroup round
end instead of the wavefront one), a last group holding padding slots, one replica, integer and
float couplings, a stop at a target energy, batches of instances too large for one workgroup
(which the batch entry point sends here), the shape of the reference's published workload
(n = 2048, R = 256, experiments/phase14_gpu.py:74-81), and a mailbox wait forced to give up from
the host.  The reference runs replicas one warp each with no exchange of energies between blocks
(cuda/ising_pt.cu:86-146); this engine has no counterpart there.
"""

import os

import numpy as np
import pytest

import oracle
from drift_amd import gpu
from drift_amd.builders.lattice import er_sparse_weighted, random_regular_gaussian
from drift_amd.builders.qubo import maxcut_ising, random_graph
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

THREADS = min(16, len(os.sched_getaffinity(0)))


def _announce(name):
    print(f"[gpu-split] {name}", flush=True)


def _check(g, o, model):
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
    ran = len(o["best_history"])   # a run stopped at a target fills the rest with its best
    assert np.array_equal(g.best_history[:ran], o["best_history"])
    assert np.all(g.best_history[ran:] == g.best_E)
    e = model.energy(g.best_s)
    assert g.best_E == pytest.approx(e, rel=1e-5, abs=1e-6)


def _run_pair(model, *, R, n_rounds, S, seed, rg, T_min=0.2, T_max=5.0, block=0, target=None,
              trace=True, cl=False):
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    beta = ladder_betas(geometric_ladder(R, T_min, T_max))
    g = gpu.parallel_tempering_gpu(
        model, n_replicas=R, n_rounds=n_rounds, sweeps_per_round=S, seed=seed, betas=beta,
        engine="split", slots_per_group=rg, block_threads=block, coloring=(cptr, cspins),
        record_history=True, return_final=True, record_trace=trace, target_energy=target,
        couplings_in_lds=cl)
    assert g.couplings_in_lds == cl
    ran = g.rounds_run if target is not None else n_rounds
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=seed, R=R, n_rounds=ran,
                              sweeps_per_round=S, beta=beta, n_threads=THREADS, trace=trace)
    return g, o, enc


@pytest.mark.parametrize("rg", [1, 2, 4, 8, 16])
def test_group_sizes_equal_oracle_float(rg):
    """A float 3-regular instance (N = 600, R = 16) with 16 / rg groups of rg slots."""
    _announce(f"float 3-regular N=600 R=16, {rg} slots per group")
    model = random_regular_gaussian(600, seed=41)
    g, o, enc = _run_pair(model, R=16, n_rounds=25, S=2, seed=7, rg=rg)
    assert enc.weight_type == 3 and g.engine == "split" and g.slots_per_group == rg
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


@pytest.mark.parametrize("R,rg", [(100, 4), (100, 8), (33, 16), (1, 1), (64, 64)])
def test_ladder_shapes_equal_oracle_int(R, rg):
    """Integer couplings (ER, weights +-1..+-3, N = 1500): R above 64 takes the workgroup round
    end; R = 100 with 8 slots per group and R = 33 with 16 leave padding slots in the last group;
    one replica has no exchange; R = 64 in one group is the LDS engine's layout on this engine."""
    _announce(f"int ER N=1500 R={R}, {rg} slots per group")
    model = er_sparse_weighted(1500, 6.0, 12)
    g, o, enc = _run_pair(model, R=R, n_rounds=12, S=2, seed=3, rg=rg, T_min=0.1, T_max=4.0)
    assert enc.weight_type == 1 and g.engine == "split"
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


def test_block_sizes_do_not_change_results():
    """The same instance at 64, 256 and 1024 threads per workgroup: identical chains."""
    _announce("block sizes")
    model = maxcut_ising(random_graph(300, p=0.05, seed=2))
    outs = []
    for block in (64, 256, 1024):
        g, o, _ = _run_pair(model, R=32, n_rounds=15, S=3, seed=11, rg=4, block=block)
        _check(g, o, model)
        assert g.block_threads == block
        outs.append(g)
    assert all(np.array_equal(outs[0].final_spins, x.final_spins) for x in outs[1:])


def test_stop_at_target_equals_oracle_run_of_that_length():
    """A run told to stop at an energy the chain reaches after a few rounds stops at the first
    round end that reaches it; the result equals the oracle's run of that many rounds."""
    _announce("stop at target")
    model = random_regular_gaussian(900, seed=5)
    full, o_full, _ = _run_pair(model, R=16, n_rounds=60, S=1, seed=2, rg=2, trace=False)
    hist = o_full["best_history"]
    target = float(hist[len(hist) // 3])
    g, o, _ = _run_pair(model, R=16, n_rounds=60, S=1, seed=2, rg=2, target=target, trace=False)
    assert 0 < g.rounds_run < 60 and g.best_E <= target + 1e-9 * max(1.0, abs(target))
    _check(g, o, model)


def test_published_workload_shape_equals_oracle():
    """The reference's published benchmark graph (maxcut_ising(random_graph(2048,
    p=min(0.1, 20/n)), seed 0), R = 256, T 0.05..5.0, experiments/phase14_gpu.py:74-81) for 10
    rounds of 4 sweeps, 256 groups of one slot: too large for one workgroup, so the split engine
    is where it runs in one launch."""
    n = 2048
    _announce("published workload shape, R=256, one slot per group")
    model = maxcut_ising(random_graph(n, p=min(0.1, 20 / n), seed=0))
    g, o, _ = _run_pair(model, R=256, n_rounds=10, S=4, seed=1, rg=1, T_min=0.05, T_max=5.0,
                        trace=False)
    _check(g, o, model)


def test_batch_of_large_instances_runs_on_the_split_engine():
    """A batch whose instances do not fit one workgroup (N 9000..9600, R = 32, float couplings)
    goes to the split engine in one persistent launch; every instance equals the oracle."""
    count, R, n_rounds, S = 20, 32, 8, 2
    _announce(f"batch of {count} large float instances on the split engine")
    models = [random_regular_gaussian(9000 + 32 * i, seed=700 + i) for i in range(count)]
    seeds = [3 + 5 * i for i in range(count)]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True)
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    for model, seed, r in zip(models, seeds, res):
        assert r.engine == "split"
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=n_rounds,
                                  sweeps_per_round=S, beta=beta, n_threads=THREADS)
        _check(r, o, model)


def test_mailbox_wait_gives_up_from_the_host(monkeypatch):
    """A wait bound of 0 ms (DRIFT_PT_BARRIER_TIMEOUT_MS, set from the host) makes the first
    mailbox wait that has to spin give up: the abort flag reaches every workgroup, every wave
    leaves, and the call fails with the engine's message.  The next call runs normally."""
    _announce("mailbox give-up")
    model = er_sparse_weighted(3000, 6.0, 4)
    monkeypatch.setenv("DRIFT_PT_BARRIER_TIMEOUT_MS", "0")
    with pytest.raises(RuntimeError, match="mailbox wait exceeded its bound"):
        gpu.parallel_tempering_gpu(model, n_replicas=32, n_rounds=40, sweeps_per_round=2, seed=6,
                                   engine="split", slots_per_group=1)
    monkeypatch.delenv("DRIFT_PT_BARRIER_TIMEOUT_MS")
    g, o, _ = _run_pair(model, R=32, n_rounds=10, S=2, seed=6, rg=1, trace=False)
    _check(g, o, model)


def test_batch_at_the_held_out_size_equals_oracle():
    """Eight random 3-regular instances of N = 65536 (the size of the learned-ladder evaluation's
    held-out instances), R = 32: 256 groups of one slot, one persistent launch, every instance
    equal to the oracle bit for bit over 6 rounds of 2 sweeps."""
    count, R, n_rounds, S = 8, 32, 6, 2
    _announce(f"batch of {count} instances of N = 65536 on the split engine")
    models = [random_regular_gaussian(65536, seed=80_000 + i) for i in range(count)]
    seeds = [17 + i for i in range(count)]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True)
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    for model, seed, r in zip(models, seeds, res):
        assert r.engine == "split" and r.slots_per_group == 1
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=n_rounds,
                                  sweeps_per_round=S, beta=beta, n_threads=THREADS)
        _check(r, o, model)


def _er_int_field(n, degree, seed):
    """Erdos-Renyi graph with integer weights in +-1..+-3 and an integer field in -2..2."""
    from drift_amd.ising import SparseIsingModel

    base = er_sparse_weighted(n, degree, seed)
    rng = np.random.default_rng(seed + 1)
    rp, ci, w = base.to_csr()
    return SparseIsingModel(rp, ci, w, rng.integers(-2, 3, size=n).astype(float))


@pytest.mark.parametrize("case", [("er", 100, 8), ("er", 16, 16), ("er_field", 32, 2),
                                  ("published", 256, 1), ("published", 256, 4)])
def test_couplings_in_lds_equal_oracle(case):
    """The split engine with every instance's couplings staged in LDS as 16-bit records (int8
    couplings, |w| <= 7, n <= 4096): ER graphs with weights +-1..+-3 (R = 100 with padding slots,
    one group of 16), an integer field, and the reference's published workload graph (degree 20,
    R = 256) -- all equal to the oracle bit for bit."""
    kind, R, rg = case
    _announce(f"couplings in LDS: {kind}, R={R}, {rg} slots per group")
    if kind == "er":
        model = er_sparse_weighted(3000, 6.0, 21)
    elif kind == "er_field":
        model = _er_int_field(2500, 9.0, 22)
    else:
        model = maxcut_ising(random_graph(2048, p=20 / 2048, seed=0))
    g, o, enc = _run_pair(model, R=R, n_rounds=10, S=3, seed=5, rg=rg, T_min=0.05, T_max=5.0,
                          cl=True)
    assert enc.weight_type == 1
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


def test_couplings_in_lds_refuse_what_does_not_fit_sixteen_bits():
    """Float couplings, or n above 4096, cannot be staged as 16-bit records: the call fails
    with the engine's message."""
    _announce("couplings in LDS refusal")
    with pytest.raises(RuntimeError, match="couplings in LDS need"):
        gpu.parallel_tempering_gpu(random_regular_gaussian(600, seed=1), n_replicas=16,
                                   n_rounds=2, engine="split", couplings_in_lds=True)
    with pytest.raises(RuntimeError, match="couplings in LDS need"):
        gpu.parallel_tempering_gpu(er_sparse_weighted(5000, 4.0, 2), n_replicas=16,
                                   n_rounds=2, engine="split", couplings_in_lds=True)


def test_mixed_batch_with_targets_equals_oracle_runs():
    """One batch holding small instances (the LDS launch) and instances too large for one
    workgroup (the split launch), each told to stop at an energy its chain reaches part-way:
    every result equals the oracle's run of the rounds it ran."""
    R, n_rounds, S = 32, 40, 1
    _announce("mixed batch with targets")
    small = [random_regular_gaussian(400 + 40 * i, seed=90_000 + i) for i in range(6)]
    large = [random_regular_gaussian(12000 + 64 * i, seed=91_000 + i) for i in range(10)]
    models = small + large
    seeds = [7 + i for i in range(len(models))]
    beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
    # targets from a first batched run: the best energy after a third of the rounds
    first = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                             n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                             record_history=True)
    targets = [float(r.best_history[n_rounds // 3]) for r in first]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.2, T_max=5.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True, targets=targets)
    assert {r.engine for r in res[:6]} == {"lds"} and {r.engine for r in res[6:]} == {"split"}
    for model, seed, t, r in zip(models, seeds, targets, res):
        assert r.rounds_run <= n_rounds // 3 + 1 and r.best_E <= t + 1e-9 * max(1.0, abs(t))
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=r.rounds_run,
                                  sweeps_per_round=S, beta=beta, n_threads=THREADS)
        _check(r, o, model)


def test_batch_autotune_with_couplings_in_lds_candidates():
    """Large int8 instances that qualify for couplings in LDS: the batch tuner times those shapes
    too, and the tuned batch equals the untuned one bit for bit."""
    count, R = 12, 64   # 64 slots of N ~ 3500 exceed one workgroup; N <= 4096 fits 16-bit records
    _announce("batch autotune with couplings-in-LDS candidates")
    gpu.autotune_clear()
    models = [er_sparse_weighted(3500 + 20 * i, 7.0, 95_000 + i) for i in range(count)]
    kw = dict(n_replicas=R, T_min=0.1, T_max=4.0, n_rounds=8, sweeps_per_round=2,
              return_final=True, slots_per_group=0)
    plain = gpu.parallel_tempering_gpu_batch(models, **kw)
    tuned = gpu.parallel_tempering_gpu_batch(models, autotune=True, **kw)
    assert plain[0].engine == tuned[0].engine == "split"
    for a, b in zip(plain, tuned):
        assert a.best_E_units == b.best_E_units and np.array_equal(a.final_spins, b.final_spins)
        assert np.array_equal(a.rung_of, b.rung_of)
