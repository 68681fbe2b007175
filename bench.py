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


def lattice_side(n: int) -> int:
    """Side L of an L^3 or L^2 lattice of n sites (0 if n is neither)."""
    for d in (3, 2):
        L = int(round(n ** (1.0 / d)))
        if L ** d == n:
            return L
    return 0


def algorithmic_bytes_per_attempt(model, enc, engine: str, replicas: int = 0) -> float:
    """Bytes one attempt touches at the declared storage format, no credit for on-chip reuse.
    Spins are int8 (b_s = 1), so own spin read + write and d neighbour reads are d + 2 bytes;
    the CSR engine keeps up to 32 replicas as one bit each of a word per position (b_s = 1/8).
    The CSR kernels add d edge records (int8 weight and 24-bit column packed in 4 bytes; int32 /
    float weight and int32 column in 8) and two int32 row pointers (plus the field if any);
    positions are relabelled by colour, so no colour list is read.  The lattice kernel computes
    neighbours from coordinates and reads one byte holding the site's bond signs."""
    d = model.nnz / model.n
    b_s = 0.125 if engine == "global" and 0 < replicas <= 32 else 1.0
    if engine in ("lattice", "slab", "pslab"):
        return b_s * (d + 1) + b_s + 1.0
    b_e = 4.0 if enc.weight.dtype.itemsize == 1 and model.n < (1 << 24) else 8.0
    b_h = 0.0 if enc.h is None else enc.h.dtype.itemsize
    return b_s * (d + 1) + b_s + d * b_e + 8.0 + b_h


def engine_format(engine: str, model, enc, replicas: int = 0) -> str:
    """Storage format the engine ran the workload in (what the algorithmic bytes count)."""
    if engine == "lattice":
        return ("int8 spins [position][replica], implicit periodic neighbours, one byte of bond "
                "signs per site")
    if engine == "slab":
        return ("int8 spins [replica][site], slabs of planes staged in LDS each sweep, implicit "
                "periodic neighbours, one byte of bond signs per site")
    if engine == "pslab":
        side = round(model.n ** (1 / 3))
        side = side if side ** 3 == model.n else round(model.n ** 0.5)
        bits = side >= 32 and side % 32 == 0
        return ("one launch for the whole run (occupancy-sized grid, bounded grid barriers); "
                "each workgroup keeps its slab resident in LDS "
                + ("as 32-site spin words and bond bit-planes (bit-sliced field), halo planes "
                   "exchanged as bits" if bits else
                   "as int8 spins and bond-sign bytes, halo planes exchanged as bytes")
                + "; int8 spins [replica][site] in HBM at the ends of the run")
    words = engine == "global" and 0 < replicas <= 32
    spins = ("spins as one 32-bit word per position (bit per replica)" if words
             else "int8 spins [position][replica]" + (" in LDS" if engine == "lds" else ""))
    rec = (f"({'4' if enc.weight.dtype.itemsize == 1 else '8'} bytes: column and "
           f"{enc.weight.dtype.name} weight)")
    if words or engine == "lds":
        return (f"{spins}; the first edge records of each row (4 or 8 on the word engine, 2-4 on "
                f"the LDS engine) in a zero-padded row-major copy read with 16-byte loads, longer "
                f"rows finished from packed CSR edge records {rec}; the algorithmic bytes count "
                f"the unpadded records")
    return f"{spins}, packed CSR edge records {rec}"


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
    traffic_file = os.path.join(ROOT, "profiles", "traffic_c4_lds.json")
    if os.path.exists(traffic_file):   # one launch runs the whole timed region
        with open(traffic_file) as f:
            roofline["traffic"] = round(json.load(f)["hbm_bytes_per_sweep"] * args.steps)
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
                   "format": engine_format(res[0].engine, models[0],
                                           gpu.encode_problem(models[0]), R),
                   "engine": res[0].engine, "wall_s": round(wall, 3),
                   "mean_best_E_per_spin": round(float((best / [m.n for m in models]).mean()), 5)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


LADDER_FAMILIES = {          # family: (T_min, T_max)
    "ferro2d": (0.8, 8.0),
    "ea3d": (0.3, 3.0),
    "rrg": (0.1, 3.0),
    "er": (0.2, 10.0),
}


def ladder_instance(family: str, seed: int):
    """One instance of the learned-ladder mix (N between 216 and 2048)."""
    from drift_amd.builders.lattice import (cubic_pmj_glass, er_sparse_weighted,
                                            random_regular_gaussian, square_ferromagnet)

    rng = np.random.default_rng(seed)
    if family == "ferro2d":
        return square_ferromagnet(int(2 * rng.integers(8, 23)))
    if family == "ea3d":
        return cubic_pmj_glass(int(rng.choice([6, 8, 10, 12])), seed)
    if family == "rrg":
        return random_regular_gaussian(int(2 * rng.integers(128, 1025)), seed)
    return er_sparse_weighted(int(rng.integers(256, 2049)), float(rng.uniform(3.0, 12.0)), seed)


def ladder_set(per_family: int, base_seed: int):
    items = []
    for fi, fam in enumerate(LADDER_FAMILIES):
        for k in range(per_family):
            items.append((fam, ladder_instance(fam, base_seed + 1000 * fi + k)))
    return items


def bench_ladder(args) -> int:
    """Learned vs geometric ladder: feedback labels -> features -> network trained by the HIP
    kernels -> held-out time-to-target, with batched vs one-instance-per-launch wall time."""
    from drift_amd import gpu, ladder

    R, budget = 32, args.steps
    t_all = time.perf_counter()
    train = ladder_set(args.train_per_family, 1)
    test = ladder_set(args.test_per_family, 777_777)
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)   # noqa: E731

    def by_family(items):
        groups = {}
        for i, (fam, m) in enumerate(items):
            groups.setdefault(fam, []).append(i)
        return groups

    # 1. feedback-optimised labels and features of the training set, one batch per family
    X, Y = np.zeros((len(train), 16)), np.zeros((len(train), R))
    t0 = time.perf_counter()
    for fam, idx in by_family(train).items():
        T_min, T_max = LADDER_FAMILIES[fam]
        models = [train[i][1] for i in idx]
        x, _ = ladder.feedback_optimise(models, T_min=T_min, T_max=T_max, R=R,
                                        iterations=5, sweeps=args.label_sweeps, damping=0.5)
        Y[idx] = x
        X[idx] = ladder.instance_features(models, T_min=T_min, T_max=T_max, R=R)
        log(f"labels {fam}: {len(idx)} instances")
    t_labels = time.perf_counter() - t0
    # 2. train the network with its own kernels
    net = ladder.LadderNet(seed=11)
    t0 = time.perf_counter()
    losses = net.fit(X, Y, epochs=args.epochs, lr=3e-3)
    t_train = time.perf_counter() - t0
    log(f"trained: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    # 3. held-out: geometric vs learned ladders, same seeds, one batched launch per family
    rows, wall_learn_batch, t_geo_batch, t_learn_batch = [], 0.0, 0.0, 0.0
    for fam, idx in by_family(test).items():
        T_min, T_max = LADDER_FAMILIES[fam]
        models = [test[i][1] for i in idx]
        xs = net.predict(ladder.instance_features(models, T_min=T_min, T_max=T_max, R=R))
        seeds = [9000 + i for i in idx]
        kw = dict(n_replicas=R, T_min=T_min, T_max=T_max, n_rounds=budget, sweeps_per_round=1,
                  seeds=seeds, record_history=True)
        t0 = time.perf_counter()
        geo = gpu.parallel_tempering_gpu_batch(models, **kw)
        t1 = time.perf_counter()
        lrn = gpu.parallel_tempering_gpu_batch(
            models, betas=[ladder.betas_from_positions(x, T_min, T_max) for x in xs], **kw)
        t2 = time.perf_counter()
        t_geo_batch += t1 - t0
        t_learn_batch += t2 - t1
        fam_rows = []
        for m, g, l in zip(models, geo, lrn):
            target = min(g.best_E, l.best_E)
            fam_rows.append(dict(family=fam, n=m.n, target=target, seed=seeds[len(fam_rows)],
                                 tt_geo=ladder.sweeps_to_target(g.best_history, target),
                                 tt_learn=ladder.sweeps_to_target(l.best_history, target),
                                 rt_geo=ladder.round_trip_rate(g, budget),
                                 rt_learn=ladder.round_trip_rate(l, budget), best_geo=g.best_E))
        # the batched learned launch stopped once its slowest instance is at the target
        slowest = max(r["tt_learn"] or budget for r in fam_rows)
        wall_learn_batch += (t2 - t1) * min(slowest, budget) / budget
        rows += fam_rows
        log(f"held-out {fam}: {len(idx)} instances")
    # 4. the reference's way: each instance alone in its own launch with the geometric ladder,
    # stopped at its own time to target (instances that never reach it run the whole budget)
    wall_geo_single, mismatch = 0.0, 0
    for r, (fam, m) in zip(rows, test):
        T_min, T_max = LADDER_FAMILIES[fam]
        t0 = time.perf_counter()
        one = gpu.parallel_tempering_gpu(m, n_replicas=R, T_min=T_min, T_max=T_max,
                                         n_rounds=budget, sweeps_per_round=1, seed=r["seed"],
                                         record_history=False)
        wall = time.perf_counter() - t0
        mismatch += int(one.best_E != r["best_geo"])       # same stream -> same answer
        wall_geo_single += wall * min(r["tt_geo"] or budget, budget) / budget
    log(f"single launches done, {mismatch} best-energy mismatches against the batched run")

    def med(v):
        return float(np.median(v)) if len(v) else None

    tt_g = np.array([r["tt_geo"] or 2 * budget for r in rows], dtype=float)
    tt_l = np.array([r["tt_learn"] or 2 * budget for r in rows], dtype=float)
    hit_g = int(sum(r["tt_geo"] is not None for r in rows))
    hit_l = int(sum(r["tt_learn"] is not None for r in rows))
    per_family = {}
    for fam in LADDER_FAMILIES:
        sel = [i for i, r in enumerate(rows) if r["family"] == fam]
        per_family[fam] = {
            "hit_geo": int(sum(rows[i]["tt_geo"] is not None for i in sel)),
            "hit_learned": int(sum(rows[i]["tt_learn"] is not None for i in sel)),
            "median_sweeps_geo": med(tt_g[sel]), "median_sweeps_learned": med(tt_l[sel]),
            "round_trip_ratio": med([rows[i]["rt_learn"] / max(rows[i]["rt_geo"], 1e-12) for i in sel]),
        }
    line = {
        "metric": "time-to-target energy, learned vs geometric ladder",
        "value": round(wall_geo_single / max(wall_learn_batch, 1e-9), 3),
        "unit": "x less wall time (batched+learned vs one-launch-per-instance+geometric)",
        "n_gpus": 1, "steps": budget, "warmup": 0, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32/f32", "data": "synthetic",
        "config": {
            "workload": "learned 32-rung ladders on held-out instances of four families "
                        "(2-D ferromagnet, 3-D +/-J glass, 3-regular N(0,1), sparse ER integer), "
                        "N 216..2048",
            "train_instances": len(train), "test_instances": len(test),
            "budget_sweeps": budget, "label_sweeps": args.label_sweeps, "epochs": args.epochs,
            "target": "best energy found by either ladder within the budget",
        },
        "held_out": {
            "reached_target_geo": hit_g, "reached_target_learned": hit_l,
            "median_sweeps_to_target_geo": med(tt_g), "median_sweeps_to_target_learned": med(tt_l),
            "sweeps_ratio_median": med(tt_g / tt_l),
            "wall_s_geo_one_per_launch": round(wall_geo_single, 3),
            "wall_s_learned_batched": round(wall_learn_batch, 3),
            "batched_s_geo_full_budget": round(t_geo_batch, 3),
            "batched_s_learned_full_budget": round(t_learn_batch, 3),
            "single_vs_batched_best_mismatches": mismatch,
            "per_family": per_family,
        },
        "timing_s": {"labels": round(t_labels, 1), "train": round(t_train, 1),
                     "total": round(time.perf_counter() - t_all, 1)},
        "train_loss": [float(losses[0]), float(losses[-1])],
    }
    print(json.dumps(line), flush=True)
    return 0


C5_FAMILIES = {            # family: (T_min, T_max)
    "ea3d": (0.3, 3.0),
    "rrg": (0.1, 3.0),
    "er": (0.2, 10.0),
}


def c5_instance(family: str, seed: int, n_target: int = 0):
    """One instance of the C5 mix: 3-D +/-J glasses, 3-regular N(0,1) graphs and sparse ER graphs
    (mean degree 3..12); N uniform on a log scale over 256..16384, or about n_target when given
    (the glass then takes the nearest cube)."""
    from drift_amd.builders.lattice import (cubic_pmj_glass, er_sparse_weighted,
                                            random_regular_gaussian)

    rng = np.random.default_rng(seed)
    n = n_target or int(round(np.exp(rng.uniform(np.log(256), np.log(16384)))))
    if family == "ea3d":
        return cubic_pmj_glass(max(6, int(round(n ** (1 / 3)))), seed)
    if family == "rrg":
        return random_regular_gaussian(2 * (n // 2), seed)
    return er_sparse_weighted(n, float(rng.uniform(3.0, 12.0)), seed)


def bench_ladder_c5(args) -> int:
    """C5: learned 32-rung ladders at scale, in two stages (each fits one call of a few minutes).
    --stage train: feedback-optimised labels for --train-per-family instances of each family (N 256..
    16384), features, network trained by the HIP kernels, saved to --net.
    --stage eval: the saved network's ladders against geometric ones on --test-per-family held-out
    instances per family at N ~ --test-n, time to target within --steps sweeps."""
    from drift_amd import ladder

    R = 32
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)   # noqa: E731
    t_all = time.perf_counter()
    if args.stage == "train":
        feats, labels, sizes, picks = [], [], [], {}
        for fi, fam in enumerate(C5_FAMILIES):
            T_min, T_max = C5_FAMILIES[fam]
            models = [c5_instance(fam, 10_000 * fi + k) for k in range(args.train_per_family)]
            t0 = time.perf_counter()
            step = 64 if args.labels == "budget" else 128   # chunks, so progress shows as it goes
            for c0 in range(0, len(models), step):
                chunk = models[c0:c0 + step]
                if args.labels == "budget":
                    x, pick = ladder.budget_labels(chunk, T_min=T_min, T_max=T_max, R=R,
                                                   budget=args.steps,
                                                   seeds=range(c0, c0 + len(chunk)))
                    picks.setdefault(fam, []).extend(int(k) for k in pick)
                else:
                    x, _ = ladder.feedback_optimise(chunk, T_min=T_min, T_max=T_max, R=R,
                                                    iterations=5, sweeps=args.label_sweeps,
                                                    damping=0.5, seeds=range(c0, c0 + len(chunk)))
                labels.append(x)
                feats.append(ladder.instance_features(chunk, T_min=T_min, T_max=T_max, R=R))
                log(f"labels {fam}: {c0 + len(chunk)}/{len(models)}, "
                    f"{time.perf_counter() - t0:.1f} s")
            sizes += [m.n for m in models]
            log(f"labels {fam}: {len(models)} instances, N {min(m.n for m in models)}.."
                f"{max(m.n for m in models)}, {time.perf_counter() - t0:.1f} s")
        X, Y = np.concatenate(feats), np.concatenate(labels)
        net = ladder.LadderNet(seed=11)
        t0 = time.perf_counter()
        losses = net.fit(X, Y, epochs=args.epochs, lr=3e-3)
        net.save(args.net)
        line = {"stage": "train", "labels": args.labels, "instances": int(X.shape[0]),
                "n_min": int(min(sizes)), "n_max": int(max(sizes)),
                "label_sweeps": args.steps if args.labels == "budget" else args.label_sweeps,
                "train_loss": [float(losses[0]), float(losses[-1])],
                "chosen_gamma": {f: {str(g): p.count(k) for k, g in enumerate(ladder.BUDGET_GAMMAS)}
                                 for f, p in picks.items()},
                "train_s": round(time.perf_counter() - t0, 1),
                "total_s": round(time.perf_counter() - t_all, 1), "net": args.net}
        print(json.dumps(line), flush=True)
        return 0

    net = ladder.LadderNet(seed=11)
    net.load(args.net)
    budget, rows = args.steps, []
    for fi, fam in enumerate(C5_FAMILIES):
        T_min, T_max = C5_FAMILIES[fam]
        models = [c5_instance(fam, 900_000 + 1000 * fi + k, n_target=args.test_n)
                  for k in range(args.test_per_family)]
        xs = net.predict(ladder.instance_features(models, T_min=T_min, T_max=T_max, R=R))
        log(f"{fam}: predicted rung 8 position {np.mean(xs[:, 8]):.3f} (geometric 0.258), "
            f"rung 16 {np.mean(xs[:, 16]):.3f} (0.516)")
        seeds = [4242 + k for k in range(len(models))]
        kw = dict(n_replicas=R, T_min=T_min, T_max=T_max, n_rounds=budget, sweeps_per_round=1,
                  seeds=seeds, record_history=True)
        t0 = time.perf_counter()
        geo = ladder.run_instances(models, **kw)
        t1 = time.perf_counter()
        lrn = ladder.run_instances(
            models, betas=[ladder.betas_from_positions(x, T_min, T_max) for x in xs], **kw)
        t2 = time.perf_counter()
        for m, g, l in zip(models, geo, lrn):
            target = min(g.best_E, l.best_E)
            rows.append(dict(family=fam, n=m.n,
                             tt_geo=ladder.sweeps_to_target(g.best_history, target),
                             tt_learn=ladder.sweeps_to_target(l.best_history, target),
                             rt_geo=ladder.round_trip_rate(g, budget),
                             rt_learn=ladder.round_trip_rate(l, budget),
                             gap=(g.best_E - l.best_E) / abs(target), engine=g.engine))
        log(f"held-out {fam}: {len(models)} instances at N {models[0].n}, geometric {t1 - t0:.1f} s, "
            f"learned {t2 - t1:.1f} s")

    def summary(sel):
        tg = np.array([r["tt_geo"] or 2 * budget for r in sel], dtype=float)
        tl = np.array([r["tt_learn"] or 2 * budget for r in sel], dtype=float)
        return {"instances": len(sel), "reached_geo": int((tg <= budget).sum()),
                "reached_learned": int((tl <= budget).sum()),
                "median_sweeps_geo": float(np.median(tg)), "median_sweeps_learned": float(np.median(tl)),
                "learned_first": int((tl < tg).sum()), "geometric_first": int((tg < tl).sum()),
                "median_rel_gap_geo_minus_learned": float(np.median([r["gap"] for r in sel])),
                "round_trip_ratio_median": float(np.median(
                    [r["rt_learn"] / max(r["rt_geo"], 1e-12) for r in sel]))}

    line = {
        "metric": "time-to-target energy, learned vs geometric ladder (held-out, large N)",
        "value": summary(rows)["reached_learned"], "unit": "held-out instances at target",
        "n_gpus": 1, "steps": budget, "warmup": 0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32/f32", "data": "synthetic",
        "config": {"workload": f"C5: learned 32-rung ladders, held-out N ~ {args.test_n} "
                               "(3-D +/-J glass at the nearest cube, 3-regular N(0,1), sparse ER)",
                   "test_per_family": args.test_per_family, "budget_sweeps": budget,
                   "target": "best energy found by either ladder within the budget",
                   "net": args.net},
        "held_out": summary(rows),
        "per_family": {f: summary([r for r in rows if r["family"] == f]) for f in C5_FAMILIES},
        "total_s": round(time.perf_counter() - t_all, 1),
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
    # the same chain on one thread, over a fifth of the budget
    sweeps1 = int(max(1, min(sweeps, budget_s / 5 / (per_sweep * cores))))
    one = oracle.pt_run_encoded(enc, (cptr, cs), seed=1, R=R, n_rounds=sweeps1, sweeps_per_round=1,
                                beta=beta, n_threads=1)
    return {
        "value": attempts / run["seconds"] / 1e9,
        "unit": "attempts/ns",
        "cores": cores,
        "kind": "port",
        "sample": f"same workload, {sweeps} sweeps+exchanges of all {R} replicas "
                  f"({attempts:.3g} attempts, {run['seconds']:.1f} s, OpenMP over replicas)",
        "value_1_thread": R * model.n * sweeps1 / one["seconds"] / 1e9,
        "sample_1_thread": f"{sweeps1} sweeps+exchanges on one thread ({one['seconds']:.1f} s)",
        "cpu_model": cpu_model(),
    }


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or '' where that is not readable."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--engine", default="auto")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--autotune", action="store_true",
                    help="with --engine auto: time every candidate engine first (the default takes "
                         "the library's engine choice)")
    ap.add_argument("--kernel-timing-steps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--instances", type=int, default=4096, help="batch size of the c4 workload")
    ap.add_argument("--train-per-family", type=int, default=128)
    ap.add_argument("--stage", default="train", choices=("train", "eval"),
                    help="ladder-c5 workload: train the network or evaluate a saved one")
    ap.add_argument("--net", default="ladder_c5_net.npz", help="ladder-c5 network file")
    ap.add_argument("--labels", default="feedback", choices=("feedback", "budget"),
                    help="ladder-c5 training labels: feedback-optimised ladders, or the best of "
                         "the ladder shapes at a budget of --steps sweeps")
    ap.add_argument("--test-n", type=int, default=65536, help="ladder-c5 held-out size")
    ap.add_argument("--test-per-family", type=int, default=16)
    ap.add_argument("--label-sweeps", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=3000)
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
    if args.workload == "ladder":
        return bench_ladder(args)
    if args.workload == "ladder-c5":
        return bench_ladder_c5(args)
    model, R, beta, desc = workload(args.workload)
    enc = gpu.encode_problem(model)
    kw = dict(n_replicas=R, betas=beta, sweeps_per_round=1, seed=1, engine=args.engine,
              block_threads=args.block, record_history=False,
              autotune=args.autotune and args.engine == "auto" and args.block == 0,
              # the bench owns the device, so the persistent engines start with an ordinary
              # dispatch (rocprofv3 cannot trace cooperative launches)
              plain_launch=True)

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
    B = algorithmic_bytes_per_attempt(model, enc, kt.engine, R)
    if kt.kernel_launches > 1:
        # The dominant kernel's average launch in the timed region: the region's time per step
        # times that kernel's share of device time, over its launches per step.  The share comes
        # from the per-launch event timing run, where every launch (sweep kernels and round ends
        # alike) is bracketed by an event pair that itself adds about 2 us.
        launches_per_step = kt.kernel_launches / args.kernel_timing_steps
        share = kt.kernel_seconds / (kt.kernel_seconds + (kt.other_kernel_seconds or 0.0))
        per_launch_s = res.seconds / args.steps * share / launches_per_step
        per_launch_attempts = kt.kernel_attempts / kt.kernel_launches
        event_us = kt.kernel_seconds / kt.kernel_launches * 1e6
    else:
        # one launch runs the whole timed region
        per_launch_s, per_launch_attempts, event_us = res.seconds, attempts, res.seconds * 1e6
    achieved = B * per_launch_attempts / per_launch_s / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": {"global": "g_color_bits" if R <= 32 else "g_color_step",
                   "lattice": "g_lattice_step",
                   "slab": "slab_sweep", "coop": "coop_run",
                   # the persistent slab engine keeps slabs as bits when L is a multiple of 32
                   "pslab": "pslab_bits_run" if lattice_side(model.n) % 32 == 0 else "pslab_run",
                   }.get(kt.engine, "pt_lds_kernel"),
        "bytes_per_attempt": round(B, 3),
        "attempts_per_launch": per_launch_attempts,
        "avg_launch_us": round(per_launch_s * 1e6, 3),
        "avg_launch_us_event_pairs": round(event_us, 3),
    }
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{kt.engine}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        if "hbm_bytes_per_launch" in tr:
            roofline["traffic"] = tr["hbm_bytes_per_launch"]
        elif "hbm_bytes_per_sweep" in tr and kt.kernel_launches <= 1:
            # one launch runs the whole timed region
            roofline["traffic"] = round(tr["hbm_bytes_per_sweep"] * args.steps)

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
            "format": engine_format(res.engine, model, enc, R),
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
