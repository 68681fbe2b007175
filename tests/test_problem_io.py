"""Problem files: the drift-csr-v1 sparse format and the reference's problem.bin / result.bin."""

import struct

import numpy as np
import pytest

from drift_amd.builders.families import pmj_spin_glass
from drift_amd.builders.lattice import er_sparse_weighted
from drift_amd.ising import SparseIsingModel
from drift_amd.problem_io import (load_problem, read_problem_bin, read_result_bin, save_problem,
                                  write_problem_bin, write_result_bin)


def _energies(model, states):
    return np.array([model.energy(s) for s in states])


def test_csr_round_trip_keeps_the_energy_function(tmp_path):
    dense = pmj_spin_glass(30, 4, field=0.3)
    info = save_problem(str(tmp_path / "g.npz"), dense, meta={"id": "pmj-30-4"})
    back = load_problem(str(tmp_path / "g.npz"))
    assert isinstance(back, SparseIsingModel) and back.n == 30
    assert back.meta["id"] == "pmj-30-4" and back.meta["fingerprint"] == info["fingerprint"]
    states = np.where(np.random.default_rng(0).random((20, 30)) < 0.5, -1.0, 1.0)
    np.testing.assert_array_equal(_energies(back, states), _energies(dense, states))


def test_csr_file_with_a_changed_coupling_is_refused(tmp_path):
    path = str(tmp_path / "g.npz")
    save_problem(path, pmj_spin_glass(12, 1))
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k].copy() for k in z.files}
    arrays["weight"][0] *= -1        # J_ij changes, J_ji does not: no longer symmetric
    np.savez(path, **arrays)
    with pytest.raises(ValueError):
        load_problem(path)


def test_large_sparse_instance_through_the_file_format(tmp_path):
    model = er_sparse_weighted(100_000, 6.0, 3)
    save_problem(str(tmp_path / "er.npz"), model)
    back = load_problem(str(tmp_path / "er.npz"))
    assert back.n == model.n and back.nnz == model.nnz
    np.testing.assert_array_equal(back.col_idx, model.col_idx)
    np.testing.assert_array_equal(back.weight, model.weight)


def test_problem_bin_layout(tmp_path):
    """Header <Q 6i 2f> then int32 row_ptr, int32 col_idx, float32 weight, float32 h, int32
    color_ptr, int32 color_spins (cuda/ising_pt.cu:40-45)."""
    model = pmj_spin_glass(10, 2, field=0.5)
    path = str(tmp_path / "problem.bin")
    write_problem_bin(path, model, n_replicas=8, n_rounds=7, sweeps_per_round=3, seed=2**40 + 5,
                      T_min=0.25, T_max=4.0)
    raw = open(path, "rb").read()
    seed, n, R, rounds, sweeps, nnz, k, tmin, tmax = struct.unpack_from("<Qiiiiiiff", raw, 0)
    assert (seed, n, R, rounds, sweeps, tmin, tmax) == (2**40 + 5, 10, 8, 7, 3, 0.25, 4.0)
    rp, ci, w = model.to_csr()
    assert nnz == ci.size and 1 <= k <= n
    assert len(raw) == 40 + 4 * (n + 1) + 4 * nnz + 4 * nnz + 4 * n + 4 * (k + 1) + 4 * n
    p = read_problem_bin(path)
    np.testing.assert_array_equal(p["row_ptr"], rp)
    np.testing.assert_array_equal(p["col_idx"], ci)
    np.testing.assert_array_equal(p["weight"], w.astype(np.float32))
    np.testing.assert_array_equal(p["h"], np.asarray(model.h, dtype=np.float32))
    # the colouring partitions 0..n-1 into independent sets
    assert sorted(p["color_spins"].tolist()) == list(range(n))
    colour = np.empty(n, int)
    for c in range(k):
        colour[p["color_spins"][p["color_ptr"][c]:p["color_ptr"][c + 1]]] = c
    for i in range(n):
        assert all(colour[j] != colour[i] for j in ci[rp[i]:rp[i + 1]])


def test_problem_bin_truncated_is_refused(tmp_path):
    path = str(tmp_path / "problem.bin")
    write_problem_bin(path, pmj_spin_glass(10, 2))
    open(path, "r+b").truncate(60)
    with pytest.raises(ValueError):
        read_problem_bin(path)


def test_result_bin_round_trip(tmp_path):
    s = np.where(np.arange(9) % 3 == 0, -1.0, 1.0)
    write_result_bin(str(tmp_path / "result.bin"), -12.5, s)
    e, back = read_result_bin(str(tmp_path / "result.bin"), 9)
    assert e == -12.5 and np.array_equal(back, s)
