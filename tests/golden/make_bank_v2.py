"""Write bank_v2.json and bank_v2_states.npz: the v2 rows (drift_amd.bank.v2_specs) with best-known
energies from independent runs of the GPU engine.

    python tests/golden/make_bank_v2.py          # needs a gfx950 device

Every row is run from the PT seeds drift_amd.bank.V2_SEEDS (four independent chains), each for
v2_rounds(spec) rounds of one sweep and one exchange on the row's ladder (drift_amd.bank.V2_RUN).
Rows whose replicas fit one workgroup's LDS run as one batched launch per family (all their rows
and seeds together; a batched instance gives what it gives alone); the others run one at a time.
The best-known energy is the lowest over the seeds; the row records every seed's energy, how many
seeds reached the best, and the provenance (the reference keeps a known energy's source beside it,
drift/benchmarks/bank.py:215-233).  The lowest state itself is stored bit-packed in
bank_v2_states.npz, so the energy can be checked on the CPU against the model."""
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from drift_amd import gpu  # noqa: E402
from drift_amd.bank import V2_RUN, V2_SEEDS, v2_model, v2_rounds, v2_specs  # noqa: E402

OUT = os.environ.get("BANK_V2_OUT", os.path.dirname(os.path.abspath(__file__)))


def fits_one_workgroup(spec):
    return V2_RUN[spec["family"]]["R"] * spec["n"] <= 150_000   # int8 spins of all replicas


def main():
    specs = v2_specs()
    runs = {s["id"]: [] for s in specs}
    t_all = time.perf_counter()
    for family in V2_RUN:
        cfg = V2_RUN[family]
        kw = dict(n_replicas=cfg["R"], T_min=cfg["T_min"], T_max=cfg["T_max"], sweeps_per_round=1,
                  record_history=False)
        small = [s for s in specs if s["family"] == family and fits_one_workgroup(s)]
        by_rounds = {}
        for s in small:
            by_rounds.setdefault(v2_rounds(s), []).append(s)
        for rounds, group in by_rounds.items():
            jobs = [(s, seed) for s in group for seed in V2_SEEDS]
            models = [v2_model(s["family"], s["size"], s["seed"]) for s, _ in jobs]
            t0 = time.perf_counter()
            res = gpu.parallel_tempering_gpu_batch(models, n_rounds=rounds,
                                                   seeds=[seed for _, seed in jobs], **kw)
            wall = time.perf_counter() - t0
            for (s, seed), r in zip(jobs, res):
                runs[s["id"]].append(dict(seed=seed, n_rounds=rounds, best_E=r.best_E,
                                          best_round=r.best_round, engine="lds (batched)",
                                          device_seconds=round(r.seconds, 3), state=r.best_s))
            print(f"[bank-v2] {family}: {len(jobs)} batched runs of {rounds} rounds, "
                  f"{wall:.1f} s", flush=True)
        for s in specs:
            if s["family"] != family or fits_one_workgroup(s):
                continue
            model = v2_model(s["family"], s["size"], s["seed"])
            for seed in V2_SEEDS:
                r = gpu.parallel_tempering_gpu(model, n_rounds=v2_rounds(s), seed=seed, **kw)
                runs[s["id"]].append(dict(seed=seed, n_rounds=v2_rounds(s), best_E=r.best_E,
                                          best_round=r.best_round, engine=r.engine,
                                          device_seconds=round(r.seconds, 3), state=r.best_s))
            print(f"[bank-v2] {s['id']}: {[x['best_E'] for x in runs[s['id']]]}", flush=True)
    rows, states = [], {}
    for s in specs:
        rs = runs[s["id"]]
        best = min(x["best_E"] for x in rs)
        win = next(x for x in rs if x["best_E"] == best)
        states[s["id"]] = np.packbits(win["state"] < 0)
        cfg = V2_RUN[s["family"]]
        rows.append(dict(
            s, best_known_energy=best,
            best_known_source=(f"gpu-pt: lowest of {len(rs)} independent seeds, "
                               f"{rs[0]['n_rounds']} rounds each"),
            seeds_reaching_best=sum(x["best_E"] == best for x in rs),
            run={"n_replicas": cfg["R"], "T_min": cfg["T_min"], "T_max": cfg["T_max"],
                 "n_rounds": rs[0]["n_rounds"], "sweeps_per_round": 1},
            runs=[{k: v for k, v in x.items() if k != "state"} for x in rs]))
    with open(os.path.join(OUT, "bank_v2.json"), "w") as f:
        json.dump({"description": "bank v2: best-known energies from independent GPU PT runs",
                   "rows": rows}, f, indent=1)
    np.savez_compressed(os.path.join(OUT, "bank_v2_states.npz"), **states)
    print(f"[bank-v2] done in {time.perf_counter() - t_all:.1f} s", flush=True)


if __name__ == "__main__":
    main()
