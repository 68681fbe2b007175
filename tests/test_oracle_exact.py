"""Exact enumeration: the oracle's Gray-code walk against the package's chunked enumerator, the
reference's bank energies and closed forms (1-D chain thermodynamics, 2-D L=2 partition function)."""

import json
import math
import os

import numpy as np
import pytest

import oracle
from drift_amd.builders.families import build_model
from drift_amd.builders.lattice import square_ferromagnet
from drift_amd.ising import IsingModel
from drift_amd.solvers.exact import all_configs, exact_ground_state, exact_thermal

BANK = os.path.join(os.path.dirname(__file__), "golden", "reference_bank_v1.json")


def _csr(model):
    rp, ci, w = model.to_csr()
    return rp, ci, w, np.asarray(model.h, dtype=np.float64)


def test_all_configs_bit_convention():
    # reference drift/solvers/exact.py:23: bit b of the index -> spin (1 - 2b)
    S = all_configs(3)
    assert S.shape == (8, 3)
    assert S[0].tolist() == [1, 1, 1]
    assert S[1].tolist() == [-1, 1, 1]
    assert S[6].tolist() == [1, -1, -1]


def test_bank_known_energies_n_le_16():
    with open(BANK) as f:
        rows = json.load(f)["instances"]
    checked = 0
    for r in rows:
        if r["known_energy"] is None or r["family"] == "tfim-chain" or r["n"] > 16:
            continue
        model, _ = build_model(r["family"], r["n"], r["seed"], r["params"])
        e_or, idx, _, _ = oracle.exact(*_csr(model))
        s, e_np, E = exact_ground_state(model)
        assert e_or == pytest.approx(r["known_energy"], abs=1e-9), r["id"]
        assert e_np == pytest.approx(r["known_energy"], abs=1e-9), r["id"]
        assert idx == int(np.argmin(E)), r["id"]
        checked += 1
    assert checked >= 50


def test_bank_fixtures_match_generators():
    with open(BANK) as f:
        fixtures = json.load(f)["fixtures"]
    assert len(fixtures) == 10
    for fx in fixtures:
        model, _ = build_model(fx["family"], fx["n"], fx["seed"], fx["params"])
        assert np.array_equal(model.J, np.asarray(fx["J"])), fx["id"]
        assert np.array_equal(model.h, np.asarray(fx["h"])), fx["id"]
        _, e, _ = exact_ground_state(IsingModel(np.asarray(fx["J"]), np.asarray(fx["h"])))
        assert e == pytest.approx(fx["known_energy"]), fx["id"]


def test_open_chain_thermal_closed_form():
    # open chain, h = 0: Z = 2 (2 cosh b)^(n-1), <E> = -(n-1) tanh(b)
    n = 12
    J = np.zeros((n, n))
    i = np.arange(n - 1)
    J[i, i + 1] = J[i + 1, i] = 1.0
    model = IsingModel(J, np.zeros(n))
    betas = np.array([0.1, 0.5, 1.0, 2.0])
    _, _, mE, mE2 = oracle.exact(*_csr(model), betas)
    ref = -(n - 1) * np.tanh(betas)
    np.testing.assert_allclose(mE, ref, rtol=1e-12, atol=1e-12)
    # variance: (n-1) sech^2(b)
    np.testing.assert_allclose(mE2 - mE ** 2, (n - 1) / np.cosh(betas) ** 2, rtol=1e-9)
    pm, _, _ = exact_thermal(model, betas)
    np.testing.assert_allclose(pm, ref, rtol=1e-12)


def test_periodic_square_L2_partition_function():
    # L=2 periodic: every neighbour pair is bonded twice, so the model is the 4-ring 0-1-3-2-0 with
    # J = 2: Z = (2 cosh 2b)^4 + (2 sinh 2b)^4 and <E> = -dlnZ/db.
    model = square_ferromagnet(2)
    for b in (0.05, 0.3, 0.9):
        c, s = 2 * np.cosh(2 * b), 2 * np.sinh(2 * b)
        Z = c ** 4 + s ** 4
        dZ = 4 * c ** 3 * 4 * np.sinh(2 * b) + 4 * s ** 3 * 4 * np.cosh(2 * b)
        e0, idx, mE, _ = oracle.exact(*_csr(model), [b])
        assert mE[0] == pytest.approx(-dZ / Z, rel=1e-12)
        assert e0 == -8.0 and idx == 0


def test_L4_square_thermal_against_numpy_enumeration():
    model = square_ferromagnet(4)
    betas = np.geomspace(0.1, 1.0, 16)
    e0, idx, mE, mE2 = oracle.exact(*_csr(model), betas)
    pm, pm2, pe0 = exact_thermal(model.to_dense(), betas)
    assert e0 == pe0 == -32.0 and idx == 0
    np.testing.assert_allclose(mE, pm, rtol=1e-10)
    np.testing.assert_allclose(mE2, pm2, rtol=1e-10)


def test_gray_code_matches_brute_force_float_field():
    rng = np.random.default_rng(3)
    n = 10
    J = rng.normal(size=(n, n))
    model = IsingModel(J, rng.normal(size=n))
    e0, idx, mE, _ = oracle.exact(*_csr(model), [0.7])
    s, e, E = exact_ground_state(model)
    assert e0 == pytest.approx(e, abs=1e-9) and idx == int(np.argmin(E))
    w = np.exp(-0.7 * (E - E.min()))
    assert mE[0] == pytest.approx((w * E).sum() / w.sum(), rel=1e-10)
