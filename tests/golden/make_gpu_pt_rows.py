"""Write gpu_pt_rows.json: the reference's published gpu-pt rows of its scale sweep (the rows its own
GPU engine produced, docs/results/scale_sweep.json), as test data for the bank harness.

    python tests/golden/make_gpu_pt_rows.py /path/to/reference

Reads the reference's JSON results file only (json.load); keeps id, family, n, seed, the published
energy and the known energy (null where the bank has none)."""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
with open(os.path.join(ref, "docs/results/scale_sweep.json")) as f:
    sweep = json.load(f)
rows = [{"id": r["id"], "family": r["family"], "n": r["n"], "seed": r["seed"],
         "published_energy": r["energy"], "known_energy": r["known_energy"]}
        for r in sweep["rows"] if r["method"] == "gpu-pt"]
out = {"source": "docs/results/scale_sweep.json (gpu-pt rows; hardware "
                 f"{sweep['hardware'].get('platform')}, n_rounds {sweep['hardware'].get('n_rounds')})",
       "rows": rows}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_pt_rows.json"), "w") as f:
    json.dump(out, f, indent=1)
print(len(rows), "rows")
