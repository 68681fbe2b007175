"""The bank rows the reference solved with its GPU engine, and the two large-n quality checks,
through this package's ``solve`` on the gfx950 engine."""

import json
import os

import pytest

from drift_amd import gpu
from drift_amd.bank import bipartite_check, crystal_check, sweep_row

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not gpu.gpu_available(), reason="no gfx950 device / engine not built"),
]

HERE = os.path.dirname(__file__)
MANIFEST = {r["id"]: r for r in json.load(open(os.path.join(HERE, "golden/reference_bank_v1.json")))["instances"]}
GPU_ROWS = json.load(open(os.path.join(HERE, "golden/gpu_pt_rows.json")))["rows"]


def test_published_gpu_pt_rows():
    """Every gpu-pt row of the reference's scale sweep (n = 20..40, 400 rounds of 4 sweeps, the row's
    seed): this engine's energy is the published one or lower, and the known optimum where the bank
    has one."""
    print(f"[bank-gpu] {len(GPU_ROWS)} rows", flush=True)
    worse = []
    for r in GPU_ROWS:
        row = sweep_row(MANIFEST[r["id"]])
        assert row.method == "gpu-pt" and row.certified is False
        if row.energy > r["published_energy"] + 1e-9:
            worse.append((r["id"], row.energy, r["published_energy"]))
        if r["known_energy"] is not None:
            assert row.energy >= r["known_energy"] - 1e-9, r["id"]
    assert not worse, worse


@pytest.mark.parametrize("L", [16, 32, 64])
def test_crystal_reaches_the_stripe_ground_state(L):
    """experiments/crystal_gpu.py: the L x L frustrated crystal reaches E/n = -2 as a defect-free
    period-4 stripe."""
    res = crystal_check(L)
    print(f"[bank-gpu] crystal {L}x{L}: {res}", flush=True)
    assert res["method"] == "gpu-pt"
    assert res["energy_per_spin"] == pytest.approx(-2.0, abs=1e-12)
    assert res["striped"] and res["period"] == 4


@pytest.mark.parametrize("n", [128, 256, 512, 1024])
def test_planted_bipartite_max_cut_is_optimal(n):
    """experiments/phase14_gpu.py:88-99: the maximum cut of a planted bipartite graph is every
    edge."""
    res = bipartite_check(n)
    print(f"[bank-gpu] bipartite: {res}", flush=True)
    assert res["optimal"], res


V2 = json.load(open(os.path.join(HERE, "golden/bank_v2.json")))["rows"]


@pytest.mark.parametrize("row", [r for r in V2 if r["seed"] == 1], ids=lambda r: r["id"])
def test_bank_v2_best_known_energy_replays_exactly(row):
    """The recorded best-known run of a v2 row replays to the recorded energy and round: the runs
    are deterministic under the stream contract, whichever engine the automatic choice takes."""
    from drift_amd.bank import v2_best_known_run

    res = v2_best_known_run(row, row["run"]["n_rounds"], seed=row["run"]["seed"])
    print(f"[bank-gpu] {row['id']}: {res.best_E} on {res.engine} in {res.seconds:.2f} s", flush=True)
    assert res.best_E == row["best_known_energy"]
    assert res.best_round == row["best_round"]
