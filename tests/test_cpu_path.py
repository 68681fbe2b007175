"""The reference's CPU-side tests for the hot-path surface, run against drift_amd:
parallel_tempering (reference tests/test_parallel_tempering.py:37-98), solve()
(tests/test_solve.py:29-94, the GPU branch through a monkeypatched engine), the dense model
(drift/ising.py) and the sparse model that carries large instances."""

import numpy as np
import pytest

from drift_amd import gpu
from drift_amd.builders.lattice import cubic_pmj_glass, er_sparse_weighted, random_regular_gaussian
from drift_amd.builders.qubo import cut_value, maxcut_ising, qubo_to_ising, random_graph
from drift_amd.ising import IsingModel, SparseIsingModel
from drift_amd.solve import solve
from drift_amd.solvers.exact import exact_ground_state
from drift_amd.solvers.parallel_tempering import PtResult, geometric_ladder, parallel_tempering


def _ferro_chain(n):
    J = np.zeros((n, n))
    for i in range(n - 1):
        J[i, i + 1] = J[i + 1, i] = 1.0
    return IsingModel(J=J, h=np.zeros(n))


def test_pt_finds_exact_maxcut():
    model = maxcut_ising(random_graph(16, p=0.5, seed=3))
    _, e_exact, _ = exact_ground_state(model)
    res = parallel_tempering(model, n_replicas=16, T_min=0.05, T_max=4.0, n_rounds=200,
                             sweeps_per_round=4, seed=1)
    assert np.isclose(res.best_E, e_exact)


def test_pt_finds_exact_spin_glass():
    rng = np.random.default_rng(7)
    n = 14
    J = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            J[i, j] = J[j, i] = rng.choice([-1.0, 1.0])
    model = IsingModel(J=J, h=rng.normal(scale=0.3, size=n))
    _, e_exact, _ = exact_ground_state(model)
    res = parallel_tempering(model, n_replicas=20, T_min=0.05, T_max=5.0, n_rounds=300,
                             sweeps_per_round=5, seed=2)
    assert np.isclose(res.best_E, e_exact)


def test_pt_ferromagnet_and_sparse_model():
    model = _ferro_chain(18)
    res = parallel_tempering(model, n_replicas=12, T_min=0.05, T_max=3.0, n_rounds=150,
                             sweeps_per_round=4, seed=0)
    assert np.isclose(res.best_E, -17)
    sp = IsingModel.from_csr(*model.to_csr(), model.h)
    res = parallel_tempering(sp, n_replicas=12, T_min=0.05, T_max=3.0, n_rounds=150,
                             sweeps_per_round=4, seed=0)
    assert np.isclose(res.best_E, -17)


def test_exchange_beats_a_single_cold_walker():
    # the reference's instance: complete +/-J on 16 spins from default_rng(11), one choice per pair
    rng = np.random.default_rng(11)
    n = 16
    J = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            J[i, j] = J[j, i] = rng.choice([-1.0, 1.0])
    model = IsingModel(J=J, h=np.zeros(n))
    _, e_exact, _ = exact_ground_state(model)
    stuck = 0
    for seed in range(4, 12):
        full = parallel_tempering(model, n_replicas=20, T_min=0.05, T_max=5.0, n_rounds=300,
                                  sweeps_per_round=5, seed=seed)
        cold = parallel_tempering(model, n_replicas=1, T_min=0.05, T_max=0.05, n_rounds=300,
                                  sweeps_per_round=5, seed=seed)
        assert np.isclose(full.best_E, e_exact)
        stuck += cold.best_E > e_exact + 1e-9
    assert stuck >= 4, "a lone cold walker should usually freeze above the ground state"


def test_ladder_is_healthy():
    temps = geometric_ladder(16, 0.1, 5.0)
    assert np.all(np.diff(temps) > 0)
    assert temps[0] == 0.1 and np.isclose(temps[-1], 5.0)
    assert geometric_ladder(1, 0.3, 9.0).tolist() == [0.3]
    res = parallel_tempering(maxcut_ising(random_graph(20, p=0.4, seed=5)), n_replicas=16,
                             n_rounds=120, seed=6)
    assert 0.0 < res.swap_rate < 1.0


def test_solve_small_is_exact_and_certified():
    model = maxcut_ising(random_graph(14, p=0.5, seed=1))
    _, e_exact, _ = exact_ground_state(model)
    sol = solve(model, exact_max=18)
    assert sol.method == "exact" and sol.certified is True
    assert np.isclose(sol.energy, e_exact) and np.isclose(model.energy(sol.s), sol.energy)


def test_solve_large_falls_back_to_cpu_and_is_not_certified():
    model = maxcut_ising(random_graph(20, p=0.5, seed=2))
    _, e_exact, _ = exact_ground_state(model, max_n=22)
    sol = solve(model, exact_max=16, use_gpu=False, n_replicas=24, n_rounds=200,
                sweeps_per_round=4)
    assert sol.method == "cpu-pt" and sol.certified is False
    assert np.isclose(model.energy(sol.s), sol.energy) and np.isclose(sol.energy, e_exact)


def test_solve_exact_bound_is_the_boundary():
    model = maxcut_ising(random_graph(12, p=0.5, seed=3))
    sol = solve(model, exact_max=12, use_gpu=False)
    assert sol.method == "exact" and sol.certified
    sol2 = solve(model, exact_max=11, use_gpu=False, n_rounds=150)
    assert sol2.method == "cpu-pt" and not sol2.certified


def test_solve_require_certified_raises():
    model = maxcut_ising(random_graph(12, p=0.5, seed=3))
    with pytest.raises(ValueError, match="certified"):
        solve(model, exact_max=8, require_certified=True, use_gpu=False)


def test_solve_gpu_dispatch_is_not_certified(monkeypatch):
    model = maxcut_ising(random_graph(12, p=0.5, seed=3))

    def fake_pt(model, **kwargs):
        s = np.ones(model.n)
        assert kwargs["n_replicas"] == 256 and kwargs["n_rounds"] == 400
        return PtResult(best_s=s, best_E=float(model.energy(s)), temperatures=np.array([0.1]),
                        swap_rate=0.0)

    monkeypatch.setattr("drift_amd.gpu.gpu_available", lambda: True)
    monkeypatch.setattr("drift_amd.gpu.parallel_tempering_gpu", fake_pt)
    sol = solve(model, exact_max=8, use_gpu=True)
    assert sol.method == "gpu-pt" and sol.certified is False


def test_dense_model_semantics():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(9, 9))
    m = IsingModel(A, rng.normal(size=9))
    assert np.allclose(m.J, m.J.T) and np.all(np.diag(m.J) == 0)
    s = rng.choice([-1.0, 1.0], size=9)
    for i in range(9):
        t = s.copy()
        t[i] = -t[i]
        assert np.isclose(m.delta_energy_flip(s, i), m.energy(t) - m.energy(s))
    S = rng.choice([-1.0, 1.0], size=(5, 9))
    assert np.allclose(m.energy_batch(S), [m.energy(x) for x in S])


def test_sparse_model_equals_dense_and_qubo_offset():
    Q = np.random.default_rng(1).integers(-5, 6, size=(10, 10)).astype(float)
    model, offset = qubo_to_ising(Q)
    sp = IsingModel.from_csr(*model.to_csr(), model.h)
    rng = np.random.default_rng(2)
    for _ in range(20):
        s = rng.choice([-1.0, 1.0], size=10)
        x = (s + 1) / 2
        assert np.isclose(sp.energy(s), model.energy(s))
        assert np.isclose(x @ (0.5 * (Q + Q.T)) @ x, model.energy(s) + offset)
        assert np.allclose(sp.local_fields(s), model.local_fields(s))
    W = random_graph(10, 0.5, seed=3)
    s = rng.choice([-1.0, 1.0], size=10)
    m = maxcut_ising(W)
    # E = -sum_{i<j} W_ij s_i s_j (h = 0) = 2 cut - total
    assert np.isclose(m.energy(s), 2 * cut_value(W, s) - np.triu(W).sum())
    with pytest.raises(ValueError, match="symmetric"):
        SparseIsingModel([0, 1, 1], [1], [1.0], np.zeros(2))


def test_sparse_generators():
    g = cubic_pmj_glass(6, 3)
    assert g.n == 216 and g.nnz == 216 * 6 and set(np.abs(g.weight)) == {1.0}
    assert np.all(np.diff(g.row_ptr) == 6)
    assert g.fingerprint() == cubic_pmj_glass(6, 3).fingerprint() != cubic_pmj_glass(6, 4).fingerprint()
    er = er_sparse_weighted(20000, 6.0, 1)
    assert abs(er.nnz / er.n - 6.0) < 0.15
    assert set(np.unique(er.weight).tolist()) <= {-3.0, -2.0, -1.0, 1.0, 2.0, 3.0}
    rr = random_regular_gaussian(512, 5)
    assert np.all(np.diff(rr.row_ptr) == 3)
    assert np.array_equal(rr.weight.astype(np.float32).astype(np.float64), rr.weight)
    d = rr.to_dense()
    assert np.allclose(d.J, d.J.T) and abs(np.std(rr.h) - 0.1) < 0.02
