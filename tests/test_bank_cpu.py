"""Bank harness on the CPU: instance fingerprints, the planted bipartite generator's draw order,
the exact branch of the runner, and exact optima for the published best-known rows small enough to
enumerate."""

import json
import os

import numpy as np
import pytest

import oracle
from drift_amd.bank import bank_model, planted_bipartite, sweep_row

HERE = os.path.dirname(__file__)
MANIFEST = {r["id"]: r for r in json.load(open(os.path.join(HERE, "golden/reference_bank_v1.json")))["instances"]}
GPU_ROWS = json.load(open(os.path.join(HERE, "golden/gpu_pt_rows.json")))["rows"]


def _loop_bipartite(n, p, seed):
    """The generator written as its pair loop: one draw per (i < n/2, j >= n/2), row-major."""
    rng = np.random.default_rng(seed)
    w = np.zeros((n, n))
    for i in range(n // 2):
        for j in range(n // 2, n):
            if rng.random() < p:
                w[i, j] = w[j, i] = 1.0
    return w


@pytest.mark.parametrize("n,p,seed", [(8, 0.5, 1), (33, 0.3, 4), (128, 0.08, 1)])
def test_planted_bipartite_draw_order(n, p, seed):
    w = planted_bipartite(n, p, seed)
    assert np.array_equal(w, _loop_bipartite(n, p, seed))
    half = n // 2
    assert not w[:half, :half].any() and not w[half:, half:].any()


def test_published_gpu_rows_are_bank_rows():
    assert len(GPU_ROWS) == 34
    for r in GPU_ROWS:
        spec = MANIFEST[r["id"]]
        assert (spec["family"], spec["n"], spec["seed"]) == (r["family"], r["n"], r["seed"])
        assert spec.get("known_energy") == r["known_energy"]
        if r["known_energy"] is not None:
            assert r["published_energy"] == r["known_energy"]


def test_exact_branch_of_the_runner_reproduces_known_energies():
    small = [s for s in MANIFEST.values()
             if s["n"] <= 12 and s["family"] != "tfim-chain" and s.get("known_energy") is not None]
    assert len(small) >= 10
    for spec in small:
        row = sweep_row(spec, use_gpu=False)
        assert row.method == "exact" and row.certified
        assert row.energy == pytest.approx(spec["known_energy"], abs=1e-9)


def test_enumeration_certifies_published_best_known_rows():
    """The reference's gpu-pt rows without a known energy at n <= 24 are optimal: exact Gray-code
    enumeration reaches the published energy and nothing lower."""
    rows = [r for r in GPU_ROWS if r["known_energy"] is None and r["n"] <= 24]
    assert len(rows) == 4
    for r in rows:
        model = bank_model(MANIFEST[r["id"]])
        rp, ci, w = model.to_csr()
        h = np.asarray(model.h, dtype=np.float64)
        e_min, _, _, _ = oracle.exact(rp, ci, w, h)
        assert e_min == pytest.approx(r["published_energy"], abs=1e-9), r["id"]


def test_bank_v2_generators_reproduce_the_recorded_fingerprints():
    from drift_amd.bank import v2_specs

    rec = {r["id"]: r for r in json.load(open(os.path.join(HERE, "golden/bank_v2.json")))["rows"]}
    specs = v2_specs()
    assert len(specs) == len(rec) == 20
    for s in specs:
        r = rec[s["id"]]
        assert (s["n"], s["nnz"], s["fingerprint"]) == (r["n"], r["nnz"], r["fingerprint"])


def test_bank_v2_best_known_energies_are_energies_of_recorded_states():
    """Each v2 best-known energy is the lowest of its independent runs and is the energy of the
    configuration stored beside it (bank_v2_states.npz), recomputed here in fp64 from the model.
    Integer rows agree exactly; float rows within 1e-10 relative (the engine tracks energies in
    2^-36 units, exact for couplings and fields of magnitude 2^-13 or more; a site whose field has
    smaller terms rounds by at most 2^-37 per flip, over 2e5 sweeps)."""
    import numpy as np

    from drift_amd.bank import V2_SEEDS, v2_model

    rows = json.load(open(os.path.join(HERE, "golden/bank_v2.json")))["rows"]
    states = np.load(os.path.join(HERE, "golden/bank_v2_states.npz"))
    for r in rows:
        runs = r["runs"]
        assert sorted(x["seed"] for x in runs) == sorted(V2_SEEDS)
        assert r["best_known_energy"] == min(x["best_E"] for x in runs)
        assert r["seeds_reaching_best"] == sum(x["best_E"] == r["best_known_energy"] for x in runs)
        model = v2_model(r["family"], r["size"], r["seed"])
        s = 1.0 - 2.0 * np.unpackbits(states[r["id"]])[: model.n].astype(np.float64)
        e = model.energy(s)
        if r["family"] == "rrg3":
            assert e == pytest.approx(r["best_known_energy"], rel=1e-10), r["id"]
        else:
            assert e == r["best_known_energy"], r["id"]
