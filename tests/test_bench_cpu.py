"""Host logic of bench.py that needs no GPU: the refusal of several GPUs, and the algorithmic
bytes per attempt the roofline figures are computed from (the per-attempt formula of the
measurement plan: own spin read and write, d neighbour spins, d edge records, two row pointers,
the field; lattices read one bond-sign byte instead of edge records)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from drift_amd import gpu  # noqa: E402
from drift_amd.builders.lattice import (cubic_pmj_glass, er_sparse_weighted,  # noqa: E402
                                        random_regular_gaussian)


def test_several_gpus_are_refused_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr
    assert out.stdout == ""


def test_lattice_bytes_per_attempt():
    # 3-D lattice, d = 6.  Bit-sliced slabs (L a multiple of 32): one bit per spin and per bond,
    # (7 + 1 + 6) / 8 = 1.75 B; byte slabs: 7 spin reads + 1 write + 1 sign byte = 9 B
    big, small = cubic_pmj_glass(32, 0), cubic_pmj_glass(8, 0)
    assert bench.declared_bytes(big, gpu.encode_problem(big), "pslab", 64)[0] == pytest.approx(1.75)
    assert bench.declared_bytes(small, gpu.encode_problem(small), "pslab", 64)[0] == pytest.approx(9.0)
    assert bench.lattice_side(small.n) == 8


def test_csr_bytes_per_attempt_by_spin_format():
    model = er_sparse_weighted(4000, 6.0, 0)
    enc = gpu.encode_problem(model)
    d = model.nnz / model.n
    # int8 weights pack with the column into 4-byte records; no field; Poisson rows longer than the
    # first chunk, so the row pointers are read
    assert enc.weight.dtype.itemsize == 1 and enc.h is None
    int8_spins, fmt8 = bench.declared_bytes(model, enc, "global", 64)
    words, fmtw = bench.declared_bytes(model, enc, "global", 32)
    assert int8_spins == pytest.approx((d + 2) + 4 * d + 8)
    # one bit of a 32-bit word per spin when R <= 32
    assert words == pytest.approx((d + 2) / 8 + 4 * d + 8)
    assert "one 32-bit word per position" in fmtw
    assert "int8 spins [position][replica]" in fmt8


def test_batched_float_bytes_per_attempt():
    # 3-regular float instances on the LDS engine: 4 spin reads + 1 write, three 8-byte records
    # (every row fits the first chunk, so no row pointers), a 4-byte field: 33 B
    model = random_regular_gaussian(200, 3)
    B, fmt = bench.declared_bytes(model, gpu.encode_problem(model), "lds", 16)
    assert B == pytest.approx(33.0)
    assert "in LDS" in fmt and "summed in double" in fmt


def test_split_engine_bytes_per_attempt_and_chunk_width():
    # the split engine counts bytes as the LDS engine does: 3-regular float rows, 33 B
    model = random_regular_gaussian(300, 4)
    B, fmt = bench.declared_bytes(model, gpu.encode_problem(model), "split", 32)
    assert B == pytest.approx(33.0)
    assert "mailboxes" in fmt
    # int8 couplings on a 3-D lattice read as CSR: every row of 6 fits a first chunk of width 6,
    # so no row pointers are read: 7 spin reads + 1 write + six 4-byte records
    lat = cubic_pmj_glass(8, 1)
    enc = gpu.encode_problem(lat)
    B6, fmt6 = bench.declared_bytes(lat, enc, "split", 32)
    assert B6 == pytest.approx(8.0 + 6 * 4.0)
    assert "first 6 edge records" in fmt6


def test_c5_batch_is_a_sub_result_with_its_own_steps():
    assert bench.SUB_DEFAULT_STEPS["c5_batch"] > 0
    assert "65536" in bench.C5_BATCH_DESC and "split engine" in bench.C5_BATCH_DESC


def test_a_failing_sub_result_keeps_the_headline(monkeypatch, capsys):
    """The default line: a sub-result that raises is recorded with its error, the later ones are
    not started, and the C4 headline still prints."""
    import argparse
    import json

    fake = {"value": 500.0, "steps": 20, "warmup": 5, "ms_per_step": 0.1, "dtype": "f32",
            "config": {"workload": "stand-in"}, "roofline": {}, "cpu_baseline": {}, "timing": {}}
    calls = []

    def single(name, *a, **kw):
        calls.append(name)
        if name == "c3":
            raise RuntimeError("GPU engine failed (status -1): stand-in")
        return dict(fake, value=1.0)

    def must_not_run(*a, **kw):
        raise AssertionError("started after a failed sub-result")

    monkeypatch.setattr(bench, "bench_c4", lambda *a, **kw: dict(fake))
    monkeypatch.setattr(bench, "bench_single", single)
    monkeypatch.setattr(bench, "bench_c5_batch", must_not_run)
    monkeypatch.setattr(bench, "bench_ttt_sample", must_not_run)
    monkeypatch.setattr(bench, "lds_probe", must_not_run)
    args = argparse.Namespace(no_cpu_baseline=True, cpu_budget=0.0, workload="all", steps=20,
                              warmup=5, repeats=1, instances=8, autotune=False, sub_steps=0,
                              plain_launch=False, no_time_to_target=False)
    assert bench.bench_throughput(args) == 0
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert line["value"] == 500.0 and calls == ["c2", "c3"]
    subs = line["sub_results"]
    assert subs["c2"]["value"] == 1.0 and "stand-in" in subs["c3"]["error"]
    assert all("skipped" in subs[k] for k in ("published", "single3", "c5_batch", "time_to_target"))


def test_time_to_target_parts_merge(tmp_path):
    """--stage eval-merge puts per-family parts of the time-to-target evaluation together: the
    summary counts every part's rows, adds their wall times, and refuses mismatched caps."""
    import json

    def row(i, fam, reached_geo, reached_lrn, rounds):
        res = lambda ok: {"rounds": rounds, "reached": ok, "best_E": -10.0 if ok else -9.0}
        return {"index": i, "family": fam, "n": 64, "target": -10.0,
                "geo_single": res(reached_geo), "learned_batched": res(reached_lrn),
                "geo_batched": res(reached_geo),
                "within_1e-4": {"geo": 5 if reached_geo else None, "learned": 5},
                "within_1e-3": {"geo": 5, "learned": 5},
                "to_geo_final": {"geo": 5, "learned": 3},
                "learned_rung8_position": 0.3, "learned_power": 0.5}
    parts = []
    for name, fam, rows, walls in (
            ("a", "ea3d", [row(0, "ea3d", True, True, 10), row(1, "ea3d", False, True, 40)],
             {"geo_single": 8.0, "learned_batched": 2.0, "geo_batched": 3.0}),
            ("b", "er", [row(134, "er", False, False, 100)],
             {"geo_single": 4.0, "learned_batched": 1.0, "geo_batched": 2.0})):
        path = tmp_path / f"{name}.json"
        path.write_text(json.dumps({"cap": 100, "families": [fam], "walls": walls, "rows": rows,
                                    "net": "net.npz", "total_s": 1.0}))
        parts.append(str(path))

    class Args:
        bestknown = os.path.join(ROOT, "tests", "golden", "c5_heldout_bestknown.json")
    Args.parts = ",".join(parts)
    line = bench.c5_eval_merge(Args)
    assert line["held_out"]["instances"] == 3
    assert line["held_out"]["reached_geo_single"] == 1 and line["held_out"]["reached_learned_batched"] == 2
    assert line["held_out"]["reached_within_half_cap_learned_batched"] == 2   # rounds 10 and 40 <= 50
    assert line["held_out"]["reached_within_quarter_cap_learned_batched"] == 1
    assert line["wall_s"]["geo_single"] == pytest.approx(12.0)
    assert line["value"] == pytest.approx(12.0 / 3.0)
    assert set(line["per_family"]) == {"ea3d", "er"}
    bad = tmp_path / "c.json"
    bad.write_text(json.dumps({"cap": 50, "families": ["rrg"], "walls": {}, "rows": [],
                               "net": "net.npz", "total_s": 0.0}))
    Args.parts = parts[0] + "," + str(bad)
    with pytest.raises(SystemExit):
        bench.c5_eval_merge(Args)


def test_time_to_target_batched_stage(tmp_path, monkeypatch):
    """--stage eval-batched reruns the batched legs over every family in one call each: leg (a)
    and its wall time come over from the part files unchanged, the batched legs' rows come from
    the calls, the learned run's sweeps to the geometric run's final energy use leg (a)'s energy,
    and the complete rows go to --out (the calls are replaced by stand-ins here)."""
    import json
    import types

    import numpy as np

    from drift_amd import gpu, ladder

    def row(i, fam, best):
        res = {"rounds": 50, "reached": False, "best_E": best}
        return {"index": i, "family": fam, "n": 0, "target": -10.0, "geo_single": res,
                "learned_batched": res, "geo_batched": res,
                "within_1e-4": {"geo": None, "learned": None}, "within_1e-3": {"geo": 7, "learned": 7},
                "to_geo_final": {"geo": 50, "learned": 40},
                "learned_rung8_position": 0.3, "learned_power": 0.5}
    parts = []
    for name, fam, rows, wall in (("a", "ea3d", [row(0, "ea3d", -9.5)], 8.0),
                                  ("b", "er", [row(134, "er", -9.0), row(135, "er", -9.9)], 4.0)):
        path = tmp_path / f"{name}.json"
        path.write_text(json.dumps({"cap": 100, "families": [fam], "rows": rows, "net": "net.npz",
                                    "walls": {"geo_single": wall, "learned_batched": 1.0,
                                              "geo_batched": 1.0}, "total_s": 1.0}))
        parts.append(str(path))
    calls = []

    def fake_batch(models, *, seeds, betas, targets, n_rounds, **kw):
        calls.append((len(models), list(seeds)))
        out = []
        for t, s in zip(targets, seeds):
            hist = np.linspace(-9.0, -10.0, 20) if s % 2 == 0 else np.full(20, -9.8)
            out.append(types.SimpleNamespace(best_E=float(hist[-1]), rounds_run=20,
                                             best_history=hist))
        return out
    monkeypatch.setattr(gpu, "parallel_tempering_gpu_batch", fake_batch)
    monkeypatch.setattr(ladder, "mix_instance", lambda fam, seed, n: types.SimpleNamespace(n=64))
    monkeypatch.setattr(ladder, "instance_features", lambda models, **kw: np.zeros((len(models), 16)))

    class FakeNet:          # the network runs on the device; a fixed ladder stands in
        def __init__(self, seed=0):
            pass

        def load(self, path):
            pass

        def predict_ladder(self, F):
            return np.array([ladder.power_positions(0.7, 32)] * len(F))
    monkeypatch.setattr(ladder, "LadderNet", FakeNet)

    class Args:
        bestknown = os.path.join(ROOT, "tests", "golden", "c5_heldout_bestknown.json")
        net = "net.npz"
        out = str(tmp_path / "out")
        slots_per_group = 0
        learned_only = False
    Args.parts = ",".join(parts)
    line = bench.c5_eval_batched(Args)
    assert [c[0] for c in calls] == [3, 3]                 # one call per ladder, all families
    assert calls[0][1] == [4242, 4242 + 134, 4242 + 135]
    assert line["wall_s"]["geo_single"] == pytest.approx(12.0)
    assert line["held_out"]["instances"] == 3
    assert line["held_out"]["reached_learned_batched"] == 2   # seeds 4242 and 4376 end at -10
    written = json.loads((tmp_path / "out" / "c5_eval_part_batched_ea3d_er_cap100.json").read_text())
    r = {x["index"]: x for x in written["rows"]}
    assert r[0]["geo_single"]["best_E"] == -9.5               # leg (a) as it was
    assert r[0]["to_geo_final"]["learned"] is not None        # reaches -9.5 on the way to -10
    assert r[135]["to_geo_final"]["learned"] is None          # -9.8 never reaches -9.9
