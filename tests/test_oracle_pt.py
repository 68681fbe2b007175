"""The oracle's parallel-tempering chain is pinned by physics, not by restating its code:

* it reaches the exact ground energies of the reference's own PT test instances
  (reference tests/test_parallel_tempering.py:37-98) and of bank rows;
* per-rung mean energies match exact Boltzmann averages within autocorrelation-aware error bars
  (2-D L=4 ferromagnet, config C1's small case), on the integer and the float path;
* per-pair exchange acceptance matches the exact value sum_{E,E'} P_r(E) P_r+1(E') min(1, e^{db dE});
* the bookkeeping is self-consistent: tracked energies equal recomputed ones, best_s has best_E,
  the best history is monotone, the rung map is a permutation.
A sign error in dE, a wrong threshold index or a reversed exchange rule fails the thermal tests.
"""

import numpy as np
import pytest

import oracle
from drift_amd.builders.families import pmj_spin_glass
from drift_amd.builders.lattice import square_ferromagnet
from drift_amd.builders.qubo import maxcut_ising, random_graph
from drift_amd.gpu import encode_problem
from drift_amd.ising import IsingModel
from drift_amd.solvers.exact import all_configs, exact_ground_state, exact_thermal
from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas


def run(model, *, R, n_rounds, S, seed, T_min=None, T_max=None, beta=None, trace=False,
        force_float=False):
    enc = encode_problem(model, force_float=force_float)
    k, cptr, cspins = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    if beta is None:
        beta = ladder_betas(geometric_ladder(R, T_min, T_max))
    out = oracle.pt_run_encoded(enc, (cptr, cspins), seed=seed, R=R, n_rounds=n_rounds,
                                sweeps_per_round=S, beta=beta, trace=trace)
    return enc, out


def test_exact_maxcut_reference_params():
    model = maxcut_ising(random_graph(16, p=0.5, seed=3))
    _, e, _ = exact_ground_state(model)
    _, out = run(model, R=16, T_min=0.05, T_max=4.0, n_rounds=200, S=4, seed=1)
    assert out["best_E"] == pytest.approx(e)


def test_exact_spin_glass_with_float_field():
    rng = np.random.default_rng(7)
    n = 14
    J = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            J[i, j] = J[j, i] = rng.choice([-1.0, 1.0])
    model = IsingModel(J=J, h=rng.normal(scale=0.3, size=n))
    _, e, _ = exact_ground_state(model)
    enc, out = run(model, R=20, T_min=0.05, T_max=5.0, n_rounds=300, S=5, seed=2)
    assert not enc.is_integer
    assert out["best_E"] == pytest.approx(e, abs=1e-5 * abs(e))


def test_ferro_chain_ground_state():
    n = 18
    J = np.zeros((n, n))
    J[np.arange(n - 1), np.arange(1, n)] = 1.0
    model = IsingModel(J + J.T, np.zeros(n))
    _, out = run(model, R=12, T_min=0.05, T_max=3.0, n_rounds=150, S=4, seed=0)
    assert out["best_E"] == -(n - 1)


def test_exchange_beats_single_cold_walker():
    model = pmj_spin_glass(16, 11)
    _, e, _ = exact_ground_state(model)
    _, full = run(model, R=20, T_min=0.05, T_max=5.0, n_rounds=300, S=5, seed=4)
    _, cold = run(model, R=1, T_min=0.05, T_max=0.05, n_rounds=300, S=5, seed=4)
    assert full["best_E"] == pytest.approx(e)
    assert cold["best_E"] >= e


def _binned_mean(x, blocks=40):
    x = np.asarray(x)
    m = x.size // blocks
    b = x[: m * blocks].reshape(blocks, m).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(blocks)


@pytest.mark.parametrize("force_float", [False, True])
def test_thermal_energies_L4_against_exact(force_float):
    model = square_ferromagnet(4)
    R = 16
    beta = ladder_betas(geometric_ladder(R, 1.0, 10.0))       # beta 1.0 .. 0.1 (config C1 range)
    enc, out = run(model, R=R, beta=beta, n_rounds=24000, S=1, seed=5, trace=True,
                   force_float=force_float)
    assert enc.is_integer != force_float
    exact_mE, _, _ = exact_thermal(model.to_dense(), beta)
    tr = out["rung_trace"][2000:]
    for r in range(R):
        m, err = _binned_mean(tr[:, r])
        assert abs(m - exact_mE[r]) < 5 * err + 0.02, (r, m, err, exact_mE[r])


def test_exchange_acceptance_matches_exact_distribution():
    model = square_ferromagnet(4)
    R = 8
    beta = ladder_betas(geometric_ladder(R, 1.2, 6.0))
    _, out = run(model, R=R, beta=beta, n_rounds=40000, S=2, seed=9)
    E = model.to_dense().energy_batch(all_configs(16))
    levels, counts = np.unique(E, return_counts=True)
    acc = out["swap_acc"] / out["swap_tried"]
    for r in range(R - 1):
        p1 = counts * np.exp(-beta[r] * (levels - levels.min()))
        p2 = counts * np.exp(-beta[r + 1] * (levels - levels.min()))
        p1 /= p1.sum()
        p2 /= p2.sum()
        d = (beta[r] - beta[r + 1]) * (levels[:, None] - levels[None, :])
        exact = float((p1[:, None] * p2[None, :] * np.minimum(1.0, np.exp(d))).sum())
        n_tried = out["swap_tried"][r]
        sigma = np.sqrt(exact * (1 - exact) / n_tried)
        assert abs(acc[r] - exact) < 6 * sigma + 0.01, (r, acc[r], exact)


@pytest.mark.parametrize("force_float", [False, True])
def test_bookkeeping_consistency(force_float):
    model = maxcut_ising(random_graph(40, p=0.3, seed=8))
    R, n_rounds = 12, 60
    enc, out = run(model, R=R, T_min=0.2, T_max=4.0, n_rounds=n_rounds, S=3, seed=3,
                   force_float=force_float)
    unit = enc.scale if enc.is_integer else 2.0 ** -36
    for a in range(R):
        e = model.energy(out["final_spins"][a].astype(np.float64))
        assert out["final_units"][a] * unit == pytest.approx(e, abs=1e-6)
    assert model.energy(out["best_s"].astype(np.float64)) == pytest.approx(out["best_E"], abs=1e-6)
    h = out["best_history"]
    assert np.all(np.diff(h) <= 0) and h[-1] == out["best_E"]
    assert sorted(out["rung_of"].tolist()) == list(range(R))
    assert np.all(out["swap_acc"] <= out["swap_tried"])
    # pairs alternate parity: even pairs tried ceil(n_rounds/2) times, odd floor(n_rounds/2)
    assert out["swap_tried"][0] == (n_rounds + 1) // 2 and out["swap_tried"][1] == n_rounds // 2
    assert np.all(out["rung_up"] + out["rung_down"] <= n_rounds)


def test_same_seed_same_chain_and_threads_do_not_matter():
    model = maxcut_ising(random_graph(30, p=0.3, seed=2))
    a = run(model, R=8, T_min=0.2, T_max=3.0, n_rounds=40, S=2, seed=77)[1]
    enc = encode_problem(model)
    k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    b = oracle.pt_run_encoded(enc, (cptr, cs), seed=77, R=8, n_rounds=40, sweeps_per_round=2,
                              beta=ladder_betas(geometric_ladder(8, 0.2, 3.0)), n_threads=1)
    assert np.array_equal(a["final_spins"], b["final_spins"])
    assert a["best_units"] == b["best_units"]
    c = run(model, R=8, T_min=0.2, T_max=3.0, n_rounds=40, S=2, seed=78)[1]
    assert not np.array_equal(a["final_spins"], c["final_spins"])


def test_fast_build_runs_the_checkers_chain():
    """The -O3 oracle build that times the CPU baseline gives the checker's chain bit for bit
    (integer and float couplings)."""
    from drift_amd import gpu
    from drift_amd.builders.lattice import er_sparse_weighted, random_regular_gaussian
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    for model in (er_sparse_weighted(300, 5.0, 2), random_regular_gaussian(200, 4)):
        enc = gpu.encode_problem(model)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        beta = ladder_betas(geometric_ladder(8, 0.2, 4.0))
        runs = [oracle.pt_run_encoded(enc, (cptr, cs), seed=3, R=8, n_rounds=40,
                                      sweeps_per_round=2, beta=beta, n_threads=2, fast=f)
                for f in (False, True)]
        assert runs[0]["best_units"] == runs[1]["best_units"]
        assert np.array_equal(runs[0]["final_spins"], runs[1]["final_spins"])
        assert np.array_equal(runs[0]["final_units"], runs[1]["final_units"])
        assert np.array_equal(runs[0]["swap_acc"], runs[1]["swap_acc"])
