"""Write bank_v2.json: the v2 rows (drift_amd.bank.v2_specs) with the best energy of their
recorded runs on the GPU engine (drift_amd.bank.v2_best_known_run, v2_rounds rounds, PT seed 1).

    python tests/golden/make_bank_v2.py          # needs a gfx950 device

The runs are deterministic, so tests replay the short ones and expect the same energies."""
import json
import os
import sys
import time

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from drift_amd.bank import V2_RUN, v2_best_known_run, v2_rounds, v2_specs  # noqa: E402

rows = []
for spec in v2_specs():
    rounds = v2_rounds(spec)
    t0 = time.perf_counter()
    res = v2_best_known_run(spec, rounds)
    wall = time.perf_counter() - t0
    cfg = V2_RUN[spec["family"]]
    row = dict(spec, best_known_energy=res.best_E, best_round=res.best_round,
               run={"n_replicas": cfg["R"], "T_min": cfg["T_min"], "T_max": cfg["T_max"],
                    "n_rounds": rounds, "sweeps_per_round": 1, "seed": 1},
               engine=res.engine, device_seconds=round(res.seconds, 4), wall_s=round(wall, 3))
    rows.append(row)
    print(json.dumps(row), flush=True)
out = os.path.join(os.environ.get("BANK_V2_OUT", os.path.dirname(os.path.abspath(__file__))), "bank_v2.json")
with open(out, "w") as f:
    json.dump({"description": "bank v2: best energies of the recorded GPU runs", "rows": rows}, f, indent=1)
