"""Throughput benchmark of the gfx950 parallel-tempering engine (one JSON line on stdout).

Workload (default): 3-D Edwards-Anderson +/-J glass, L = 32 periodic (N = 32768), 64 replicas on a
geometric beta ladder 0.2 .. 3.0, one sweep then one exchange attempt per step.  A step is one
sweep of all 64 replicas (64 x 32768 spin-flip attempts) plus the device-side exchange.

Three device runs: a warm-up (--warmup steps), the timed run (--steps steps; device events around
the sweep + exchange loop, inputs already in HBM), and a short run that times every sweep-kernel
launch with device events for the roofline figure.  The CPU baseline is the oracle
(oracle/pt_oracle.c, the same chain) on a bounded sample of the same workload on this host's cores.

    python bench.py [--steps K] [--warmup W] [--workload c2|c1|c3] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def workload(name: str):
    from drift_amd.builders.lattice import cubic_pmj_glass, er_sparse_weighted, square_ferromagnet
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    if name == "c2":
        model = cubic_pmj_glass(32, 0)
        R, beta = 64, ladder_betas(geometric_ladder(64, 1.0 / 3.0, 5.0))
        desc = ("3-D Edwards-Anderson +/-J glass, L=32 periodic (N=32768), R=64, beta 0.2..3.0 "
                "geometric, exchange every sweep")
    elif name == "c1":
        model = square_ferromagnet(16)
        R, beta = 16, ladder_betas(geometric_ladder(16, 1.0, 10.0))
        desc = "2-D ferromagnet, L=16 periodic (N=256), R=16, beta 0.1..1.0, exchange every sweep"
    elif name == "c4":
        return None, 16, None, ("4096 random 3-regular instances, N uniform over even values in "
                                "[64, 2048], J ~ N(0,1) and h ~ N(0,0.1^2) in float32, R=16 each, "
                                "beta 0.2..5.0 geometric, all in one persistent batched launch")
    elif name == "c3":
        model = er_sparse_weighted(100000, 6.0, 0)
        R, beta = 32, ladder_betas(geometric_ladder(32, 0.2, 20.0))
        desc = ("Erdos-Renyi N=100000 mean degree 6, weights in {+-1,+-2,+-3}, R=32, "
                "beta 0.05..5.0, exchange every sweep")
    else:
        raise SystemExit(f"unknown workload {name}")
    return model, R, beta, desc


def algorithmic_bytes_per_attempt(model, enc, engine: str) -> float:
    """Bytes one attempt touches at the declared storage format, no credit for on-chip reuse.
    Spins are int8, so own spin read + write and d neighbour reads are d + 2 bytes.  The CSR
    kernels add d (weight + int32 column) and two int32 row pointers (plus the field if any);
    positions are relabelled by colour, so no colour list is read.  The lattice kernel computes
    neighbours from coordinates and reads one byte holding the site's bond signs."""
    d = model.nnz / model.n
    b_s = 1.0
    if engine in ("lattice", "slab"):
        return b_s * (d + 1) + b_s + 1.0
    b_w = enc.weight.dtype.itemsize
    b_h = 0.0 if enc.h is None else enc.h.dtype.itemsize
    return b_s * (d + 1) + b_s + d * (b_w + 4.0) + 8.0 + b_h


def copy_calibration(nbytes: int) -> int:
    """One device-to-device copy of nbytes (torch's copy kernel) for scaling rocprof's
    FETCH_SIZE / WRITE_SIZE counters against a transfer of known size."""
    import torch

    n = nbytes // 4
    src = torch.arange(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4}), flush=True)
    return 0


def c4_instances(count: int, seed: int = 0):
    """The batched workload: `count` random 3-regular float instances of uneven size."""
    from drift_amd.builders.lattice import random_regular_gaussian

    rng = np.random.default_rng(seed)
    sizes = 2 * rng.integers(32, 1025, size=count)
    return [random_regular_gaussian(int(n), seed=10_000 + i) for i, n in enumerate(sizes)]


def bench_c4(args) -> int:
    """Batched uneven instances: whole-job attempts/ns of one persistent launch."""
    from drift_amd import gpu
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    count = args.instances
    print(f"[bench] building {count} instances", file=sys.stderr, flush=True)
    models = c4_instances(count)
    R = 16
    kw = dict(n_replicas=R, T_min=0.2, T_max=5.0, sweeps_per_round=1)
    print(f"[bench] warm-up {args.warmup} steps", file=sys.stderr, flush=True)
    gpu.parallel_tempering_gpu_batch(models, n_rounds=args.warmup, **kw)
    print(f"[bench] timed {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res = gpu.parallel_tempering_gpu_batch(models, n_rounds=args.steps, **kw)
    wall = time.perf_counter() - t0
    total_n = sum(m.n for m in models)
    attempts = R * total_n * args.steps
    secs = res[0].seconds
    value = attempts / secs / 1e9
    d = 3.0
    B = (d + 1) + 1 + d * (4 + 4) + 8 + 4          # int8 spins, fp32 J + int32 column, row ptrs, fp32 h
    roofline = {"bound": "hbm", "achieved": round(B * attempts / secs / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(B * attempts / secs / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": None, "kernel": "pt_lds_kernel", "bytes_per_attempt": B,
                "attempts_per_launch": attempts, "avg_launch_us": round(secs * 1e6, 1)}
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        cores = len(os.sched_getaffinity(0))
        sample = models[:64]
        beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
        tot_s, tot_a = 0.0, 0
        for m in sample:
            enc = gpu.encode_problem(m)
            k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
            o = oracle.pt_run_encoded(enc, (cptr, cs), seed=1, R=R, n_rounds=50, sweeps_per_round=1,
                                      beta=beta, n_threads=cores)
            tot_s += o["seconds"]
            tot_a += R * m.n * 50
        cpu = {"value": tot_a / tot_s / 1e9, "unit": "attempts/ns", "cores": cores, "kind": "port",
               "sample": f"first 64 instances of the batch, 50 sweeps each, one at a time "
                         f"(OpenMP over replicas; {tot_a:.3g} attempts, {tot_s:.1f} s)"}
    best = np.array([r.best_E for r in res])
    line = {
        "metric": "spin-flips/ns on 1 GPU", "value": round(value, 3), "unit": "attempts/ns",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload("c4")[3], "instances": count, "total_spins": int(total_n),
                   "engine": res[0].engine, "wall_s": round(wall, 3),
                   "mean_best_E_per_spin": round(float((best / [m.n for m in models]).mean()), 5)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(model, enc, R, beta, budget_s: float):
    import oracle

    cores = len(os.sched_getaffinity(0))
    k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    probe = oracle.pt_run_encoded(enc, (cptr, cs), seed=1, R=R, n_rounds=2, sweeps_per_round=1,
                                  beta=beta, n_threads=cores)
    per_sweep = max(probe["seconds"] / 2, 1e-6)
    sweeps = int(max(2, min(10000, budget_s / per_sweep)))
    run = oracle.pt_run_encoded(enc, (cptr, cs), seed=1, R=R, n_rounds=sweeps, sweeps_per_round=1,
                                beta=beta, n_threads=cores)
    attempts = R * model.n * sweeps
    return {
        "value": attempts / run["seconds"] / 1e9,
        "unit": "attempts/ns",
        "cores": cores,
        "kind": "port",
        "sample": f"same workload, {sweeps} sweeps+exchanges of all {R} replicas "
                  f"({attempts:.3g} attempts, {run['seconds']:.1f} s, OpenMP over replicas)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--engine", default="auto")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--kernel-timing-steps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--instances", type=int, default=4096, help="batch size of the c4 workload")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many bytes (rocprof counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this engine has no "
              "multi-GPU path", file=sys.stderr)
        return 2

    from drift_amd import gpu

    if not gpu.gpu_available():
        print("bench.py: GPU engine not available (no gfx950 device or library not built)",
              file=sys.stderr)
        return 1
    if args.workload == "c4":
        return bench_c4(args)
    model, R, beta, desc = workload(args.workload)
    enc = gpu.encode_problem(model)
    kw = dict(n_replicas=R, betas=beta, sweeps_per_round=1, seed=1, engine=args.engine,
              block_threads=args.block, record_history=False)

    print(f"[bench] warm-up {args.warmup} steps", file=sys.stderr, flush=True)
    gpu.parallel_tempering_gpu(model, n_rounds=args.warmup, **kw)
    print(f"[bench] timed {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res = gpu.parallel_tempering_gpu(model, n_rounds=args.steps, **kw)
    wall = time.perf_counter() - t0
    attempts = R * model.n * args.steps
    value = attempts / res.seconds / 1e9
    print(f"[bench] {value:.2f} attempts/ns over {res.seconds:.3f} s (wall {wall:.3f} s), "
          f"engine {res.engine}", file=sys.stderr, flush=True)

    print("[bench] kernel timing run", file=sys.stderr, flush=True)
    kt = gpu.parallel_tempering_gpu(model, n_rounds=args.kernel_timing_steps, time_kernels=True, **kw)
    B = algorithmic_bytes_per_attempt(model, enc, kt.engine)
    per_launch_s = kt.kernel_seconds / max(kt.kernel_launches, 1)
    per_launch_attempts = kt.kernel_attempts / max(kt.kernel_launches, 1)
    achieved = B * per_launch_attempts / per_launch_s / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": {"global": "g_color_step", "lattice": "g_lattice_step",
                   "slab": "slab_sweep"}.get(kt.engine, "pt_lds_kernel"),
        "bytes_per_attempt": round(B, 3),
        "attempts_per_launch": per_launch_attempts,
        "avg_launch_us": round(per_launch_s * 1e6, 3),
    }
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{kt.engine}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(model, enc, R, beta, args.cpu_budget)

    line = {
        "metric": "spin-flips/ns on 1 GPU",
        "value": round(value, 3),
        "unit": "attempts/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res.seconds * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32" if enc.is_integer else "f32",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "format": ("int8 spins [position][replica], implicit periodic neighbours, one "
                       "byte of bond signs per site" if res.engine == "lattice" else
                       "int8 spins [replica][site], slabs of planes staged in LDS, implicit "
                       "periodic neighbours, one byte of bond signs per site"
                       if res.engine == "slab" else
                       f"int8 spins [position][replica], CSR int32 columns, "
                       f"{enc.weight.dtype.name} weights"),
            "engine": res.engine,
            "n": model.n,
            "replicas": R,
            "sweeps_per_step": 1,
            "attempts_per_step": R * model.n,
            "wall_s": round(wall, 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
