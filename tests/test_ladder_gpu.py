"""The ladder network's HIP kernels against the float64 oracle, training, and the feedback ladder
on the GPU engine."""

import numpy as np
import pytest

from drift_amd import gpu, ladder
from drift_amd.builders.lattice import square_ferromagnet
from oracle import mlp_oracle as M

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]


def _announce(name):
    print(f"[ladder-gpu] {name}", flush=True)


def _targets(F):
    """A smooth family of monotone ladders controlled by the first two features."""
    j = np.linspace(0.0, 1.0, 32)
    a = 1.0 + 0.6 * np.tanh(F[:, :1])
    b = 0.15 * np.tanh(F[:, 1:2])
    x = j[None, :] ** a + b * np.sin(np.pi * j)[None, :]
    x = np.maximum.accumulate(np.clip(x, 0, 1), axis=1)
    x[:, 0], x[:, -1] = 0.0, 1.0
    return x


def test_forward_matches_oracle():
    _announce("forward vs oracle")
    net = ladder.LadderNet(seed=3)
    X = np.random.default_rng(0).normal(size=(37, 16)).astype(np.float32)
    got = net.raw_forward(X)
    ref = M.forward(net.params.astype(np.float64), X.astype(np.float64), bf16=True)
    np.testing.assert_allclose(got, ref, atol=2e-5)
    assert np.all(np.diff(got, axis=1) > 0) and np.all(got[:, 0] == 0)


def test_gradients_match_oracle():
    _announce("gradients vs oracle")
    net = ladder.LadderNet(seed=4)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(64, 16)).astype(np.float32)
    T = _targets(X).astype(np.float32)
    g, loss = net.raw_gradients(X, T)
    p = net.params.astype(np.float64)
    ref = M.gradients(p, X.astype(np.float64), T.astype(np.float64), bf16=True)
    assert loss == pytest.approx(M.loss(p, X.astype(np.float64), T.astype(np.float64)), rel=1e-4)
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err
    assert np.max(np.abs(g - ref)) < 1e-4 * np.max(np.abs(ref)) + 1e-9


def test_training_fits_a_smooth_ladder_family():
    _announce("training")
    rng = np.random.default_rng(2)
    F = rng.normal(size=(1024, 16))
    T = _targets(F)
    net = ladder.LadderNet(seed=5)
    losses = net.fit(F, T, epochs=1500, lr=3e-3)
    assert losses[-1] < losses[0] / 20, (losses[0], losses[-1])
    Fh = rng.normal(size=(256, 16))
    Th = _targets(Fh)
    held = np.mean((net.predict(Fh) - Th)[:, 1:-1] ** 2)
    mean_ladder = np.mean((T.mean(axis=0)[None, :] - Th)[:, 1:-1] ** 2)
    assert held < 0.15 * mean_ladder, (held, mean_ladder)


def test_params_round_trip_and_seed_determinism():
    _announce("params")
    a, b, c = ladder.LadderNet(seed=9), ladder.LadderNet(seed=9), ladder.LadderNet(seed=10)
    assert np.array_equal(a.params, b.params) and not np.array_equal(a.params, c.params)
    p = np.random.default_rng(3).normal(size=ladder.N_PARAMS).astype(np.float32)
    a.params = p
    assert np.array_equal(a.params, p)
    with pytest.raises(RuntimeError):
        a.raw_forward(np.zeros((0, 16), dtype=np.float32))


def test_feedback_ladder_increases_round_trips_on_ferromagnets():
    _announce("feedback ladder")
    models = [square_ferromagnet(L) for L in (16, 24, 32, 16, 24, 32)]
    T_min, T_max = 0.8, 8.0
    x, _ = ladder.feedback_optimise(models, T_min=T_min, T_max=T_max, iterations=4, sweeps=500)
    assert np.all(x[:, 0] == 0) and np.allclose(x[:, -1], 1) and np.all(np.diff(x, axis=1) >= 0)
    sweeps = 6000
    geo = gpu.parallel_tempering_gpu_batch(models, n_replicas=32, T_min=T_min, T_max=T_max,
                                           n_rounds=sweeps, sweeps_per_round=1, seeds=range(50, 56))
    opt = gpu.parallel_tempering_gpu_batch(
        models, n_replicas=32, T_min=T_min, T_max=T_max, n_rounds=sweeps, sweeps_per_round=1,
        seeds=range(50, 56), betas=[ladder.betas_from_positions(xi, T_min, T_max) for xi in x])
    rg = sum(ladder.round_trip_rate(r, sweeps) for r in geo)
    ro = sum(ladder.round_trip_rate(r, sweeps) for r in opt)
    print(f"[ladder-gpu] round trips per replica-sweep: geometric {rg:.3e}, feedback {ro:.3e}")
    assert ro > 1.2 * rg


def test_learned_ladder_reaches_target_more_often_on_held_out_glasses():
    """The whole pipeline at test scale: feedback labels on 48 random 3-regular N(0,1) glasses,
    the network trained by its HIP kernels, then 16 held-out instances run with the geometric and
    the predicted ladders on the same streams.  The target of each instance is the lower of the
    two best energies; the learned ladder must reach it far more often and usually first."""
    from drift_amd.builders.lattice import random_regular_gaussian

    _announce("learned vs geometric, held out")
    T_min, T_max, budget = 0.1, 3.0, 10000

    def instance(seed):
        rng = np.random.default_rng(seed)
        return random_regular_gaussian(int(2 * rng.integers(128, 385)), seed)

    train = [instance(100 + i) for i in range(48)]
    test = [instance(5000 + i) for i in range(16)]
    x, _ = ladder.feedback_optimise(train, T_min=T_min, T_max=T_max, iterations=5, sweeps=1000)
    net = ladder.LadderNet(seed=5)
    losses = net.fit(ladder.instance_features(train, T_min=T_min, T_max=T_max), x, epochs=1500)
    assert losses[-1] < 1e-2 * losses[0]
    xs = net.predict(ladder.instance_features(test, T_min=T_min, T_max=T_max))
    assert np.all(np.diff(xs, axis=1) > 0)
    kw = dict(n_replicas=32, T_min=T_min, T_max=T_max, n_rounds=budget, sweeps_per_round=1,
              seeds=range(70, 86), record_history=True)
    geo = gpu.parallel_tempering_gpu_batch(test, **kw)
    lrn = gpu.parallel_tempering_gpu_batch(
        test, betas=[ladder.betas_from_positions(v, T_min, T_max) for v in xs], **kw)
    tg, tl = [], []
    for g, l in zip(geo, lrn):
        target = min(g.best_E, l.best_E)
        tg.append(ladder.sweeps_to_target(g.best_history, target) or 2 * budget)
        tl.append(ladder.sweeps_to_target(l.best_history, target) or 2 * budget)
    tg, tl = np.array(tg), np.array(tl)
    hit_g, hit_l = int((tg <= budget).sum()), int((tl <= budget).sum())
    print(f"[ladder-gpu] reached target: geometric {hit_g}/16, learned {hit_l}/16; "
          f"median sweeps {np.median(tg):.0f} vs {np.median(tl):.0f}")
    assert hit_l >= hit_g + 6
    assert int((tl < tg).sum()) >= 10
    assert np.median(tl) < np.median(tg)


def _net_c5():
    import os

    net = ladder.LadderNet(seed=0)
    net.load(os.path.join(os.path.dirname(ladder.__file__), "ladder_net_c5.npz"))
    return net


def test_committed_network_on_held_out_instances_at_full_size():
    """The committed network (drift_amd/ladder_net_c5.npz) on one held-out instance of each family
    at N ~ 65536: ladders monotone with the family's ends fixed, a batched run equal to the same
    instance run alone, and reported energies equal to the energies of the reported states."""
    _announce("committed network, held-out N ~ 65536, one per family")
    net = _net_c5()
    held = ladder.heldout_instances(67, 65536)
    for k in (0, 67, 134):
        fam, model = held[k]
        T_min, T_max = ladder.MIX_FAMILIES[fam]
        x = net.predict_ladder(ladder.instance_features([model], T_min=T_min, T_max=T_max))[0]
        b = ladder.betas_from_positions(x, T_min, T_max)
        assert np.all(np.diff(b) < 0)
        assert b[0] == pytest.approx(1 / T_min) and b[-1] == pytest.approx(1 / T_max)
        kw = dict(n_replicas=32, T_min=T_min, T_max=T_max, n_rounds=150, sweeps_per_round=1,
                  return_final=True)
        batched = gpu.parallel_tempering_gpu_batch([model, model], seeds=[7, 8], betas=[b, b], **kw)
        alone = gpu.parallel_tempering_gpu(model, seed=7, betas=b, **kw)
        assert alone.best_E_units == batched[0].best_E_units
        assert np.array_equal(alone.final_spins, batched[0].final_spins)
        assert alone.best_E == pytest.approx(model.energy(alone.best_s), rel=1e-10)
        e64 = model.energy_batch(alone.final_spins.astype(np.float64))
        assert np.allclose(alone.final_E * alone.best_E / alone.best_E_units, e64, rtol=1e-10)
        print(f"[ladder-gpu] {fam} N={model.n}: learned power "
              f"{np.log(x[8]) / np.log(8 / 31):.2f}, best {alone.best_E:.2f}", flush=True)


def test_committed_network_gains_on_held_out_random_graphs():
    """The committed network's ladders against the geometric ones on six held-out 3-regular
    graphs at N ~ 65536, 3000 sweeps on the same streams: the learned ladders end lower."""
    _announce("committed network vs geometric, six held-out 3-regular graphs")
    net = _net_c5()
    held = [m for f, m in ladder.heldout_instances(6, 65536) if f == "rrg"]
    T_min, T_max = ladder.MIX_FAMILIES["rrg"]
    xs = net.predict_ladder(ladder.instance_features(held, T_min=T_min, T_max=T_max))
    kw = dict(n_replicas=32, T_min=T_min, T_max=T_max, n_rounds=3000, sweeps_per_round=1,
              seeds=range(60, 66))
    geo = gpu.parallel_tempering_gpu_batch(held, **kw)
    lrn = gpu.parallel_tempering_gpu_batch(
        held, betas=[ladder.betas_from_positions(x, T_min, T_max) for x in xs], **kw)
    lower = sum(l.best_E < g.best_E for l, g in zip(lrn, geo))
    gaps = [(g.best_E - l.best_E) / abs(g.best_E) for l, g in zip(lrn, geo)]
    print(f"[ladder-gpu] learned lower on {lower}/6; relative gains {np.round(gaps, 5)}", flush=True)
    assert lower >= 5
