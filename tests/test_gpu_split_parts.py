"""GPU parity of the split engine with replica groups split into parts (parts_per_group > 1).

Workgroup (group g, part q) holds group g's whole state but sweeps only part q's share of each
colour class; after every colour step the parts of a group trade that class's new spins as
step-tagged bit words, and after every round their energy changes.  The chain is the one every
engine runs, so each case is compared with the CPU oracle bit for bit (best energy units, round
and rung, best state, final spins and energies, rung permutation, swap and round-trip counters,
history and trace).

Colour classes of at most two positions per thread are swept whole by every part instead of
traded; the small instances here run 64-thread workgroups so that most of their classes are
traded, and one case at 128 threads trades the three largest classes of seven.  Cases: float,
int8 and int32 couplings; parts counts that do and do not divide a class's words; groups of one
slot (the packed spin-word path) and of several (the strided path); R above 64 (the workgroup
round end); padding slots in the last group; one replica; groups x parts x R of 1024; a stop at
a target; the Erdos-Renyi workload of N = 10^5 and R = 32 on 256 workgroups; a batch whose
groups fill the chip only in parts; a caller colouring with empty classes; the default engine's
choice of parts; 70 000 rounds past the round tags' wrap, against the per-launch engine; the
refusals; and a wait forced to give up from the host.  The reference runs each replica in one
warp with no exchange between blocks (cuda/ising_pt.cu:86-146); this engine has no counterpart
there.
"""

import os

import numpy as np
import pytest

import oracle
from drift_amd import gpu
from drift_amd.builders.lattice import er_sparse_weighted, random_regular_gaussian
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

THREADS = min(16, len(os.sched_getaffinity(0)))


@pytest.fixture(autouse=True)
def _short_wait_bound(monkeypatch):
    # a wait that cannot end fails the call after 5 s instead of 20
    monkeypatch.setenv("DRIFT_PT_BARRIER_TIMEOUT_MS", "5000")


def _announce(name):
    print(f"[gpu-split-parts] {name}", flush=True)


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
    ran = len(o["best_history"])
    assert np.array_equal(g.best_history[:ran], o["best_history"])
    assert np.all(g.best_history[ran:] == g.best_E)
    e = model.energy(g.best_s)
    assert g.best_E == pytest.approx(e, rel=1e-5, abs=1e-6)


def _run_pair(model, *, R, n_rounds, S, seed, rg, parts, T_min=0.2, T_max=5.0, target=None,
              trace=True, block=64, coloring=None):
    enc = gpu.encode_problem(model)
    if coloring is None:
        k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    else:
        cptr, cspins = coloring
    beta = ladder_betas(geometric_ladder(R, T_min, T_max))
    g = gpu.parallel_tempering_gpu(
        model, n_replicas=R, n_rounds=n_rounds, sweeps_per_round=S, seed=seed, betas=beta,
        engine="split", slots_per_group=rg, parts_per_group=parts, block_threads=block,
        coloring=(cptr, cspins),
        record_history=True, return_final=True, record_trace=trace, target_energy=target)
    assert g.engine == "split" and g.slots_per_group == rg and g.parts_per_group == parts
    ran = g.rounds_run if target is not None else n_rounds
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=seed, R=R, n_rounds=ran,
                              sweeps_per_round=S, beta=beta, n_threads=THREADS, trace=trace)
    return g, o, enc


@pytest.mark.parametrize("rg,parts", [(4, 2), (4, 3), (4, 5), (2, 4), (1, 4)])
def test_float_parts_equal_oracle(rg, parts):
    """A float 3-regular instance (N = 600, R = 16): parts counts that divide the class words or
    not, groups of several slots (strided spin words) and of one (packed spin words)."""
    _announce(f"float 3-regular N=600 R=16, {rg} slots x {parts} parts")
    model = random_regular_gaussian(600, seed=41)
    g, o, enc = _run_pair(model, R=16, n_rounds=20, S=2, seed=7, rg=rg, parts=parts)
    assert enc.weight_type == 3
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


def test_caller_colouring_with_empty_classes_in_parts():
    """Empty colour classes first, in the middle and last: parts skip them in the sweep, the
    trade and the round's energy ride (which goes with the last traded class)."""
    _announce("float 3-regular N=600 R=16, 4 slots x 4 parts, empty classes")
    model = random_regular_gaussian(600, seed=41)
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    cptr = [int(x) for x in cptr]
    m = k // 2
    cptr = [0] + cptr[: m + 1] + [cptr[m]] + cptr[m + 1:] + [cptr[-1]]
    g, o, _ = _run_pair(model, R=16, n_rounds=20, S=2, seed=7, rg=4, parts=4,
                        coloring=(np.array(cptr), cspins))
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


@pytest.mark.parametrize("R,rg,parts,block", [(100, 8, 2, 64), (33, 16, 3, 64), (1, 1, 4, 64),
                                              (32, 1, 8, 64), (32, 1, 8, 128)])
def test_int_parts_equal_oracle(R, rg, parts, block):
    """Integer couplings (ER, weights +-1..+-3, N = 1500): R = 100 takes the workgroup round end
    and leaves padding slots in the last group, as R = 33 does; one replica has no exchange;
    32 one-slot groups in 8 parts (256 workgroups) each read the other 31 slots' energies from
    the parts of their own index; at 128 threads the three largest classes are traded and the
    others swept by every part."""
    _announce(f"int ER N=1500 R={R}, {rg} slots x {parts} parts, {block} threads")
    model = er_sparse_weighted(1500, 6.0, 12)
    g, o, enc = _run_pair(model, R=R, n_rounds=10, S=2, seed=3, rg=rg, parts=parts, T_min=0.1,
                          T_max=4.0, block=block)
    assert enc.weight_type == 1
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


def test_int32_couplings_in_parts_equal_oracle():
    """Integer couplings too wide for int8 (ER weights x 97 with a field, so int32 records and
    thresholds), one-slot groups in 4 parts at 128 threads: the int32 kernel with parts."""
    from drift_amd.ising import SparseIsingModel

    _announce("parts: int32 couplings")
    base = er_sparse_weighted(3000, 5.0, 7)
    rng = np.random.default_rng(5)
    model = SparseIsingModel(base.row_ptr, base.col_idx, base.weight * 97.0,
                             rng.integers(-60, 61, size=base.n).astype(np.float64))
    g, o, enc = _run_pair(model, R=16, n_rounds=8, S=2, seed=13, rg=1, parts=4, T_min=20.0,
                          T_max=600.0, block=128)
    assert enc.weight_type == 2
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


def test_parts_stop_at_target():
    """Stop at an energy the chain reaches after a few rounds: equals the oracle's run of that
    many rounds."""
    _announce("parts: stop at target")
    model = random_regular_gaussian(900, seed=5)
    full, o_full, _ = _run_pair(model, R=16, n_rounds=40, S=1, seed=2, rg=2, parts=3, trace=False)
    hist = o_full["best_history"]
    target = float(hist[len(hist) // 3])
    g, o, _ = _run_pair(model, R=16, n_rounds=40, S=1, seed=2, rg=2, parts=3, target=target,
                        trace=False)
    assert 0 < g.rounds_run < 40 and g.best_E <= target + 1e-9 * abs(target)
    _check(g, o, model)


def test_erdos_renyi_1e5_on_256_workgroups():
    """The Erdos-Renyi workload (N = 10^5, mean degree 6, weights +-1..+-3, R = 32) as 32 groups
    of one slot in 8 parts: 256 workgroups; the five largest of its seven classes are traded as
    spin words, the two smallest swept by every part."""
    _announce("ER N=1e5 R=32, 1 slot x 8 parts")
    model = er_sparse_weighted(100_000, 6.0, 1)
    g, o, _ = _run_pair(model, R=32, n_rounds=4, S=1, seed=5, rg=1, parts=8, T_min=0.2,
                        T_max=5.0, block=1024)
    _check(g, o, model)
    assert np.array_equal(g.rung_energy_trace, o["rung_trace"])


def test_batch_with_parts_equals_oracle():
    """A batch of three instances too large for one workgroup (ER, N = 6000, R = 32) whose 32
    one-slot groups would leave the chip idle; with 4 parts per group their 384 workgroups fill it,
    so the batch takes one persistent split launch, and every instance equals the oracle."""
    _announce("parts: batch of three ER N=6000")
    R, n_rounds, S = 32, 6, 2
    models = [er_sparse_weighted(6000, 6.0, 30 + i) for i in range(3)]
    seeds = [9 + i for i in range(3)]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.1, T_max=4.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True,
                                           slots_per_group=1, parts_per_group=4,
                                           block_threads=256)
    beta = ladder_betas(geometric_ladder(R, 0.1, 4.0))
    for model, seed, r in zip(models, seeds, res):
        assert r.engine == "split" and r.parts_per_group == 4 and r.slots_per_group == 1
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=n_rounds,
                                  sweeps_per_round=S, beta=beta, n_threads=THREADS)
        _check(r, o, model)


def test_batch_mixing_one_part_and_four_equals_oracle():
    """A batch whose instances ask for different parts counts (1, 4 and 4: 288 workgroups, so the
    batch takes the split launch) runs in one launch of the kernel with parts; the instance in one
    part gathers its energies through the part mailboxes too.  Every instance equals the
    oracle."""
    _announce("parts: batch mixing 1, 4 and 4 parts")
    R, n_rounds, S = 32, 5, 2
    models = [er_sparse_weighted(6000, 6.0, 40 + i) for i in range(3)]
    seeds = [19 + i for i in range(3)]
    res = gpu.parallel_tempering_gpu_batch(models, n_replicas=R, T_min=0.1, T_max=4.0,
                                           n_rounds=n_rounds, sweeps_per_round=S, seeds=seeds,
                                           return_final=True, record_history=True,
                                           slots_per_group=1, parts_per_group=[1, 4, 4],
                                           block_threads=256)
    beta = ladder_betas(geometric_ladder(R, 0.1, 4.0))
    for model, seed, r, parts in zip(models, seeds, res, (1, 4, 4)):
        assert r.engine == "split" and r.parts_per_group == parts
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=seed, R=R, n_rounds=n_rounds,
                                  sweeps_per_round=S, beta=beta, n_threads=THREADS)
        _check(r, o, model)


def test_default_engine_takes_parts_for_few_colour_classes():
    """Without tuning, one instance of n >= 8000 with at most four colour classes and fewer
    replicas than compute units runs in one-slot groups in parts (a 3-regular graph of N = 8192 at
    R = 16: 16 parts, 1024 threads), and equals the oracle; an Erdos-Renyi graph (five or more
    classes) keeps the per-launch engine."""
    _announce("default engine: parts for few colour classes")
    model = random_regular_gaussian(8192, seed=8)
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    assert k <= 4
    beta = ladder_betas(geometric_ladder(16, 0.2, 5.0))
    g = gpu.parallel_tempering_gpu(model, n_replicas=16, n_rounds=6, sweeps_per_round=1, seed=21,
                                   betas=beta, coloring=(cptr, cspins), record_history=True,
                                   return_final=True)
    assert g.engine == "split" and g.slots_per_group == 1 and g.parts_per_group == 16
    assert g.block_threads == 1024
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=21, R=16, n_rounds=6, sweeps_per_round=1,
                              beta=beta, n_threads=THREADS)
    _check(g, o, model)
    er = er_sparse_weighted(16384, 3.0, 2)
    assert gpu.parallel_tempering_gpu(er, n_replicas=16, n_rounds=2).engine == "global"


@pytest.mark.parametrize("shape", [(1, 16, 1024), (4, 1, 256)])
def test_split_runs_past_the_round_tag_wrap(shape):
    """70 000 rounds, past the 2^16 wrap of the mailboxes' round tags, of a 3-regular N = 8192,
    R = 16 instance: one-slot groups in 16 parts (the default's shape) and groups of 4 slots
    without parts, each against the per-launch engine on the same chain -- best record, final
    spins and energies, rung permutation, swap and round-trip counters equal."""
    rg, parts, block = shape
    _announce(f"split {rg} slots x {parts} parts: 70000 rounds")
    model = random_regular_gaussian(8192, seed=8)
    beta = ladder_betas(geometric_ladder(16, 0.2, 5.0))
    kw = dict(n_replicas=16, n_rounds=70_000, sweeps_per_round=1, seed=3, betas=beta,
              record_history=False, return_final=True)
    g = gpu.parallel_tempering_gpu(model, engine="split", slots_per_group=rg,
                                   parts_per_group=parts, block_threads=block, **kw)
    ref = gpu.parallel_tempering_gpu(model, engine="global", **kw)
    assert g.engine == "split" and ref.engine == "global"
    assert (g.best_E_units, g.best_round, g.best_rung) == (ref.best_E_units, ref.best_round,
                                                           ref.best_rung)
    assert np.array_equal(g.best_s, ref.best_s)
    assert np.array_equal(g.final_spins, ref.final_spins)
    assert np.array_equal(g.final_E, ref.final_E)
    assert np.array_equal(g.rung_of, ref.rung_of)
    assert np.array_equal(g.swap_accepted, ref.swap_accepted)
    assert np.array_equal(g.round_trips, ref.round_trips)


def test_parts_refusals():
    """More parts than the largest colour class has 32-position words, and parts combined with
    couplings in LDS, are refused with a message."""
    _announce("parts: refusals")
    model = random_regular_gaussian(200, seed=3)    # classes of well under 8 words
    with pytest.raises(RuntimeError, match="32-position words of the largest colour class"):
        gpu.parallel_tempering_gpu(model, n_replicas=4, n_rounds=2, engine="split",
                                   slots_per_group=1, parts_per_group=64)
    small = er_sparse_weighted(1200, 4.0, 2)
    with pytest.raises(RuntimeError, match="do not combine with parts"):
        gpu.parallel_tempering_gpu(small, n_replicas=4, n_rounds=2, engine="split",
                                   slots_per_group=1, parts_per_group=2, couplings_in_lds=True)


def test_parts_wait_gives_up_from_the_host(monkeypatch):
    """A wait bound of 0 ms makes the first wait that has to spin give up: the abort flag reaches
    every workgroup, every wave leaves, and the call fails with its message."""
    _announce("parts: give-up")
    model = er_sparse_weighted(3000, 6.0, 4)
    monkeypatch.setenv("DRIFT_PT_BARRIER_TIMEOUT_MS", "0")
    with pytest.raises(RuntimeError, match="a mailbox wait exceeded its bound"):
        gpu.parallel_tempering_gpu(model, n_replicas=16, n_rounds=50, engine="split",
                                   slots_per_group=1, parts_per_group=8, block_threads=64)
    monkeypatch.setenv("DRIFT_PT_BARRIER_TIMEOUT_MS", "5000")
    g = gpu.parallel_tempering_gpu(model, n_replicas=16, n_rounds=5, engine="split",
                                   slots_per_group=1, parts_per_group=8, block_threads=64)
    assert g.parts_per_group == 8
