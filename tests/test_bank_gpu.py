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


@pytest.mark.parametrize("row_id", ["v2.ea3d.L000008.s001", "v2.ea3d.L000032.s002",
                                    "v2.er-sparse.n001000.s001", "v2.er-sparse.n010000.s002",
                                    "v2.rrg3.n000512.s001", "v2.rrg3.n008192.s002"])
def test_bank_v2_rows_match_the_oracle(row_id):
    """Short runs of v2 rows on the row's own ladder, through the engine the automatic choice
    takes at that size (LDS, persistent slabs, per-launch CSR), equal the CPU oracle bit for bit."""
    import numpy as np

    import oracle
    from drift_amd.bank import V2_RUN, V2_SEEDS, v2_best_known_run, v2_model
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    row = next(r for r in V2 if r["id"] == row_id)
    cfg = V2_RUN[row["family"]]
    rounds = 12 if row["n"] > 20000 else 40
    print(f"[bank-gpu] {row_id}: {rounds} rounds against the oracle", flush=True)
    res = v2_best_known_run(row, rounds, seed=V2_SEEDS[0])
    model = v2_model(row["family"], row["size"], row["seed"])
    enc = gpu.encode_problem(model)
    k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    o = oracle.pt_run_encoded(enc, (cptr, cs), seed=V2_SEEDS[0], R=cfg["R"], n_rounds=rounds,
                              sweeps_per_round=1,
                              beta=ladder_betas(geometric_ladder(cfg["R"], cfg["T_min"], cfg["T_max"])),
                              n_threads=min(16, len(os.sched_getaffinity(0))))
    assert res.best_E_units == o["best_units"]
    assert (res.best_round, res.best_rung) == (o["best_round"], o["best_rung"])
    assert np.array_equal(res.best_s.astype(np.int8), o["best_s"])
    assert np.array_equal(res.rung_of, o["rung_of"])
