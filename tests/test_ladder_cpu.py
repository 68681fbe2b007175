"""Ladder mathematics and the network oracle on the CPU.

* feedback update: a ladder whose up-fraction f is already linear in the rung index is a fixed
  point; end points stay put; the output is monotone; a kink in f pulls rungs towards it.
* position <-> temperature maps, and the geometric ladder is x_j = j/(R-1) (the reference's
  drift/solvers/parallel_tempering.py:29-37).
* the float64 network oracle: analytic gradients against central finite differences (which pins
  the oracle the GPU kernels are compared with), the monotone head, and BF16 rounding values.
"""

import numpy as np
import pytest

from drift_amd import ladder
from drift_amd.solvers.parallel_tempering import geometric_ladder
from oracle import mlp_oracle as M


def test_geometric_positions_are_the_reference_ladder():
    T = geometric_ladder(32, 0.2, 5.0)
    x = ladder.positions_from_temperatures(T)
    np.testing.assert_allclose(x, np.linspace(0, 1, 32), atol=1e-12)
    np.testing.assert_allclose(ladder.temperatures_from_positions(x, 0.2, 5.0), T, rtol=1e-12)


def test_feedback_fixed_point_and_invariants():
    R = 32
    x = np.sort(np.random.default_rng(0).random(R))
    x[0], x[-1] = 0.0, 1.0
    f_lin = np.linspace(1.0, 0.0, R)
    np.testing.assert_allclose(ladder.feedback_update(x, f_lin, damping=0.0), x, atol=1e-12)
    f = np.concatenate([np.linspace(1, 0.9, 10), np.linspace(0.5, 0.0, R - 10)])   # drop at rung 9-10
    new = ladder.feedback_update(np.linspace(0, 1, R), f, damping=0.0)
    assert new[0] == 0.0 and new[-1] == 1.0 and np.all(np.diff(new) >= 0)
    gaps = np.diff(new)
    assert gaps[9] < np.median(gaps)          # rungs gather where f falls steeply


def test_up_fraction_pins_and_fills():
    up = np.array([10, 8, 0, 3, 5, 0], dtype=float)
    down = np.array([0, 2, 0, 7, 2, 9], dtype=float)
    f = ladder.up_fraction(up, down)
    assert f[0] == 1.0 and f[-1] == 0.0 and np.all(np.diff(f) <= 0)
    assert 0.3 <= f[2] <= 0.8                 # interpolated across the unvisited rung


def test_time_to_target():
    h = np.array([-1.0, -3.0, -3.0, -7.0, -7.0])
    assert ladder.sweeps_to_target(h, -7.0) == 4
    assert ladder.sweeps_to_target(h, -8.0) is None


def test_mlp_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(scale=0.15, size=M.NPARAMS)
    X = rng.normal(size=(6, 16))
    T = np.sort(rng.random((6, 32)), axis=1)
    T[:, 0], T[:, -1] = 0.0, 1.0
    g = M.gradients(p, X, T, bf16=False)
    h = 1e-6
    for i in rng.choice(M.NPARAMS, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(pp, X, T, False) - M.loss(pm, X, T, False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3), i


def test_mlp_head_is_monotone_with_fixed_ends():
    rng = np.random.default_rng(2)
    p = rng.normal(scale=1.0, size=M.NPARAMS)
    x = M.forward(p, rng.normal(size=(50, 16)) * 5)
    assert np.all(x[:, 0] == 0.0) and np.allclose(x[:, -1], 1.0)
    assert np.all(np.diff(x, axis=1) > 0)


def test_bf16_rounding_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625], dtype=np.float32)
    r = M.bf16_round(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins; 1 + 3*2^-9 rounds up
    assert r.tolist() == [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625]


def test_feature_vector_has_sixteen_entries():
    assert ladder.N_FEATURES == 16 and M.IN == 16 and M.NPARAMS == ladder.N_PARAMS == 39328


def test_power_shapes_and_budget_label_choice(monkeypatch):
    """gamma = 1 is the geometric ladder, gamma > 1 moves rungs towards the cold end; the label of
    an instance is the shape with the lowest best energy, the earliest best round on ties."""
    T = geometric_ladder(32, 0.1, 5.0)
    assert np.allclose(ladder.power_positions(1.0), ladder.positions_from_temperatures(T))
    assert np.allclose(1.0 / ladder.betas_from_positions(ladder.power_positions(1.0), 0.1, 5.0), T)
    assert ladder.power_positions(2.0)[8] < ladder.power_positions(1.0)[8] < ladder.power_positions(0.5)[8]

    class Res:
        def __init__(self, E, rnd):
            self.best_E, self.best_round = E, rnd

    # per shape (gammas 0.5, 0.7, 1, 1.4, 2) and instance: (best energy, round it was first seen)
    table = {0.5: [(-10, 50), (-7, 9), (-3, 1)], 0.7: [(-10, 40), (-8, 9), (-3, 1)],
             1.0: [(-9, 1), (-8, 5), (-3, 1)], 1.4: [(-10, 45), (-8, 7), (-3, 1)],
             2.0: [(-10, 60), (-6, 1), (-3, 1)]}
    calls = []

    def fake_run(models, *, betas, **kw):
        x = ladder.positions_from_temperatures(1.0 / np.asarray(betas[0]))
        g = min(table, key=lambda gam: np.abs(ladder.power_positions(gam) - x).max())
        calls.append(g)
        return [Res(*table[g][i]) for i in range(len(models))]

    monkeypatch.setattr(ladder, "run_instances", fake_run)
    x, pick = ladder.budget_labels([object()] * 3, T_min=0.1, T_max=5.0, budget=100)
    assert calls == list(ladder.BUDGET_GAMMAS)
    # instance 0: -10 reached first by gamma 0.7 (round 40); instance 1: -8 first by gamma 1.0
    # (round 5); instance 2: all equal, the first shape wins
    assert [ladder.BUDGET_GAMMAS[k] for k in pick] == [0.7, 1.0, 0.5]
    assert np.allclose(x[1], ladder.power_positions(1.0))


def test_heldout_best_known_energies_are_energies_of_recorded_states():
    """The held-out instances' best-known energies (tests/golden/c5_heldout_bestknown.json) are the
    lowest of their four recorded runs and the energies of the states stored beside them,
    recomputed in fp64 from the regenerated instances."""
    import json
    import os

    from drift_amd import ladder

    here = os.path.join(os.path.dirname(__file__), "golden")
    bk = json.load(open(os.path.join(here, "c5_heldout_bestknown.json")))
    states = np.load(os.path.join(here, "c5_heldout_states.npz"))
    held = ladder.heldout_instances(bk["per_family"], bk["test_n"])
    assert len(bk["rows"]) == len(held) == 3 * bk["per_family"]
    for r in bk["rows"][::7]:   # every seventh row: about 30 instances of N ~ 65536
        fam, model = held[r["index"]]
        assert fam == r["family"] and model.n == r["n"]
        assert len(r["runs"]) == 4
        assert r["best_known_energy"] == min(x["best_E"] for x in r["runs"])
        s = 1.0 - 2.0 * np.unpackbits(states[str(r["index"])])[: model.n].astype(np.float64)
        e = model.energy(s)
        assert e == pytest.approx(r["best_known_energy"], rel=1e-10), r["index"]


def test_committed_network_is_weights_only_and_predicts_monotone_ladders():
    """drift_amd/ladder_net_c5.npz holds plain arrays (39,328 parameters and the feature scaler)
    and no pickled objects."""
    import os

    from drift_amd import ladder

    path = os.path.join(os.path.dirname(ladder.__file__), "ladder_net_c5.npz")
    with np.load(path, allow_pickle=False) as d:
        assert d["params"].shape == (ladder.N_PARAMS,) and d["params"].dtype == np.float32
        for k in ("mu", "sd"):
            assert d[k].shape == (ladder.N_FEATURES,)
        assert sorted(d.files) == ["mu", "params", "sd"]
    # the projection onto the labelled family keeps the ends and is monotone
    x = ladder.nearest_power_positions(np.linspace(0, 1, 32) ** 0.3)
    assert x[0] == 0.0 and x[-1] == 1.0 and np.all(np.diff(x) > 0)
