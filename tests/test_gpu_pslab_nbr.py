"""GPU parity of the bit-sliced persistent slab engine with neighbour-scoped synchronisation.

Runs of the pslab engine on lattices whose side is a multiple of 32 take the neighbour-synchronised
kernel unless they record the best energy of every round or the rung trace: halos wait only for
the adjacent slabs of the same replica, exchange partners meet in round-tagged mailboxes of their
two rungs, and each slot keeps its own best record, the run's best being their lexicographic
minimum.  These tests hold it to the oracle bit for bit -- best energy, round, rung and state,
final spins and energies, the rung permutation, swap, round-trip and rung-occupation counters --
over 2-D and 3-D lattices, replica counts that are not multiples of 8 (the padding slots take no
part), one slot per rung pair up to 100 replicas, slabs of one plane (small R) to eight, one to
three sweeps per round, and one to eight waves per workgroup; then against the grid-barrier kernel
on the full C2 lattice for 2*10^4 sweeps.
"""

import os

import numpy as np
import pytest

import oracle
from drift_amd import gpu
from drift_amd.builders.lattice import cubic_pmj_glass, square_ferromagnet
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

THREADS = min(16, len(os.sched_getaffinity(0)))


def _announce(name):
    print(f"[gpu-pslab-nbr] {name}", flush=True)


def _run_and_compare(model, *, R, n_rounds, S, seed, beta, block=0):
    enc = gpu.encode_problem(model)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    g = gpu.parallel_tempering_gpu(model, n_replicas=R, n_rounds=n_rounds, sweeps_per_round=S,
                                   seed=seed, betas=beta, engine="pslab", block_threads=block,
                                   coloring=(cptr, cspins), return_final=True)
    o = oracle.pt_run_encoded(enc, (cptr, cspins), seed=seed, R=R, n_rounds=n_rounds,
                              sweeps_per_round=S, beta=beta, n_threads=THREADS)
    assert g.engine == "pslab"
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
    assert g.best_E == pytest.approx(model.energy(g.best_s), abs=1e-9)
    return g


@pytest.mark.parametrize("case", [(3, 32, 64, 1), (3, 32, 16, 2), (3, 32, 5, 3), (3, 32, 2, 1),
                                  (3, 32, 1, 2), (3, 32, 80, 1), (3, 32, 100, 2), (2, 64, 12, 2),
                                  (2, 32, 48, 1), (3, 64, 8, 1)])
def test_neighbour_sync_equals_oracle(case):
    """(dimension, side, replicas, sweeps per round)."""
    dim, L, R, S = case
    _announce(f"dim={dim} L={L} R={R} S={S}")
    model = square_ferromagnet(L) if dim == 2 else cubic_pmj_glass(L, 7 + R)
    beta = ladder_betas(geometric_ladder(R, 0.3, 4.0))
    _run_and_compare(model, R=R, n_rounds=40, S=S, seed=L + R + S, beta=beta)


@pytest.mark.parametrize("block", [64, 128, 512])
def test_neighbour_sync_workgroup_sizes(block):
    """One wave (the round end and the halo waits in the only wave, nothing drawn ahead), two and
    eight waves (the other waves draw the next sweep's uniforms while wave 0 waits)."""
    _announce(f"block={block}")
    beta = ladder_betas(geometric_ladder(16, 0.3, 4.0))
    _run_and_compare(cubic_pmj_glass(32, 11), R=16, n_rounds=30, S=2, seed=block, beta=beta,
                     block=block)


def test_neighbour_sync_c2_first_thousand_sweeps():
    """Config C2 (3-D +/-J, L=32, R=64, beta 3.0 .. 0.2) for 10^3 sweeps against the oracle."""
    _announce("C2 L=32 R=64 1000 sweeps")
    beta = ladder_betas(geometric_ladder(64, 1.0 / 3.0, 5.0))
    _run_and_compare(cubic_pmj_glass(32, 0), R=64, n_rounds=1000, S=1, seed=1, beta=beta)


def test_neighbour_sync_equals_grid_barrier_over_a_long_run(monkeypatch):
    """The neighbour-synchronised and grid-barrier kernels on the full C2 lattice for 2*10^4
    sweeps, far past what the oracle runs in a test: identical chains."""
    _announce("C2 lattice, 20000 sweeps, neighbour vs grid synchronisation")
    model = cubic_pmj_glass(32, 0)
    beta = ladder_betas(geometric_ladder(64, 1.0 / 3.0, 5.0))
    fp = {}
    for sync in ("grid", "neighbour"):
        monkeypatch.setenv("DRIFT_PT_PSLAB_SYNC", sync)
        r = gpu.parallel_tempering_gpu(model, n_replicas=64, betas=beta, n_rounds=20000,
                                       sweeps_per_round=1, seed=5, engine="pslab",
                                       return_final=True)
        assert r.engine == "pslab"
        fp[sync] = (int(r.best_E_units), int(r.best_round), int(r.best_rung),
                    r.best_s.astype(np.int8).tobytes(), r.final_spins.tobytes(),
                    np.asarray(r.final_E).tobytes(), np.asarray(r.rung_of).tobytes(),
                    np.asarray(r.swap_accepted).tobytes(), np.asarray(r.round_trips).tobytes(),
                    np.asarray(r.rung_up).tobytes())
    assert fp["grid"] == fp["neighbour"]
