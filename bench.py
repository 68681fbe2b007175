"""Throughput benchmark of the gfx950 parallel-tempering engine (one JSON line on stdout).

Headline (default): config C4 -- 4096 independent random 3-regular instances, N uniform over the
even values of [64, 2048], float32 J ~ N(0,1) and h ~ N(0, 0.1^2), 16 replicas each on a geometric
beta ladder 0.2 .. 5.0, all in one persistent batched launch.  A step is one sweep of every replica
of every instance plus the device-side exchange (69.7 M spin-flip attempts).  The initial states are
drawn by an untimed launch of their own; each timed run is the launch that runs --steps rounds, timed
by device events with the inputs already in HBM.  The value is the median of --repeats timed runs.

Sub-results, each with its own roofline and CPU baseline, ride on the same line:
  c2         3-D Edwards-Anderson +/-J glass, L = 32 (N = 32768), R = 64, beta 0.2 .. 3.0
  c3         Erdos-Renyi N = 10^5, mean degree 6, weights in {+-1, +-2, +-3}, R = 32, beta 0.05 .. 5.0
  published  the reference's own timed workload (experiments/phase14_gpu.py:74-81): max-cut of
             random_graph(2048, p = min(0.1, 20/n), seed 0), R = 256, T 0.05 .. 5, rounds of four
             sweeps; its only published rate is 1.29 attempts/ns on an RTX 5060 Ti
             (docs/results/PHASE14-results.md:140-142), a different card: context, not a target

Rooflines count the bytes one attempt touches at the storage format the engine declares (no
credit for on-chip reuse); the fixed per-attempt figure of the measurement plan (8.75 B for C2,
50 B for C3, 45 B for C4) is shown beside it.  Kernels whose state lives in LDS are also put
against this device's LDS read rate, measured by a probe kernel at the start of the run.  What
limits each kernel is taken from its rocprofv3 counters (profiles/limiters.json).

    python bench.py [--steps K] [--warmup W] [--repeats 5] [--workload all|c4|c2|c3|c1|published]
                    [--no-cpu-baseline] [--plain-launch]

--plain-launch starts the persistent engines (pslab, coop) with an ordinary dispatch, for profiler
commands; timed runs use the cooperative launch.  The ladder workloads (--workload ladder,
ladder-c5) measure time to target instead.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
PROFILES = os.path.join(ROOT, "profiles")
PUBLISHED_RTX5060TI = 1.29   # attempts/ns, docs/results/PHASE14-results.md:140-142

# the fixed per-attempt byte figures of the measurement plan, per config
SURVEY_BYTES = {"c2": 8.75, "c3": 50.0, "c4": 45.0}


def workload(name: str):
    """(model, R, betas, sweeps per round, description) of a single-instance workload."""
    from drift_amd.builders.lattice import cubic_pmj_glass, er_sparse_weighted, square_ferromagnet
    from drift_amd.builders.qubo import maxcut_ising, random_graph
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    if name == "c2":
        return (cubic_pmj_glass(32, 0), 64, ladder_betas(geometric_ladder(64, 1.0 / 3.0, 5.0)), 1,
                "3-D Edwards-Anderson +/-J glass, L=32 periodic (N=32768), R=64, beta 0.2..3.0 "
                "geometric, one sweep and one exchange per step")
    if name == "c1":
        return (square_ferromagnet(16), 16, ladder_betas(geometric_ladder(16, 1.0, 10.0)), 1,
                "2-D ferromagnet, L=16 periodic (N=256), R=16, beta 0.1..1.0, one sweep and one "
                "exchange per step")
    if name == "c3":
        return (er_sparse_weighted(100000, 6.0, 0), 32, ladder_betas(geometric_ladder(32, 0.2, 20.0)),
                1, "Erdos-Renyi N=100000 mean degree 6, weights in {+-1,+-2,+-3}, R=32, "
                   "beta 0.05..5.0, one sweep and one exchange per step")
    if name == "published":
        n = 2048
        return (maxcut_ising(random_graph(n, p=min(0.1, 20.0 / n), seed=0)), 256,
                ladder_betas(geometric_ladder(256, 0.05, 5.0)), 4,
                "the reference's timed workload: max-cut of random_graph(2048, p=min(0.1, 20/n), "
                "seed=0), R=256, T 0.05..5.0, four sweeps and one exchange per step")
    raise SystemExit(f"unknown workload {name}")


C4_DESC = ("4096 random 3-regular instances, N uniform over even values in [64, 2048], J ~ N(0,1) "
           "and h ~ N(0,0.1^2) in float32, R=16 each, beta 0.2..5.0 geometric, all in one "
           "persistent batched launch; one sweep and one exchange of every instance per step")


def c4_instances(count: int, seed: int = 0):
    """The batched workload: `count` random 3-regular float instances of uneven size."""
    from drift_amd.builders.lattice import random_regular_gaussian

    rng = np.random.default_rng(seed)
    sizes = 2 * rng.integers(32, 1025, size=count)
    return [random_regular_gaussian(int(n), seed=10_000 + i) for i, n in enumerate(sizes)]


def lattice_side(n: int) -> int:
    """Side L of an L^3 or L^2 lattice of n sites (0 if n is neither)."""
    for d in (3, 2):
        L = int(round(n ** (1.0 / d)))
        if L ** d == n:
            return L
    return 0


def kernel_name(engine: str, model, R: int) -> str:
    return {"global": "g_color_bits" if R <= 32 else "g_color_step",
            "lattice": "g_lattice_step", "slab": "slab_sweep", "coop": "coop_run",
            "pslab": "pslab_bits_run" if lattice_side(model.n) % 32 == 0 else "pslab_run",
            }.get(engine, "pt_lds_kernel")


def declared_bytes(model, enc, engine: str, R: int) -> tuple[float, str]:
    """Bytes one attempt touches at the storage format the engine declares, no credit for on-chip
    reuse: B = b_s (d + 1) + b_s + d (b_J + b_idx) + b_ptr + b_list + b_h, with b_s the bytes of a
    spin, b_J of a coupling, b_idx of a column, b_ptr the row pointers (read only when some row is
    longer than the padded first chunk of edge records), b_list 0 (positions are renumbered by
    colour, so no colour list is read) and b_h the field.  Returns (B, format)."""
    d = enc.col_idx.size / model.n
    if engine == "pslab" and lattice_side(model.n) % 32 == 0:
        # bit-sliced slabs: one bit per spin and per bond, neighbours implicit
        return (d + 2) / 8.0 + d / 8.0, (
            "one launch for the whole run; each workgroup keeps its slab of one replica in LDS as "
            "32-site spin words and bond bit-planes (one bit per spin, one bit per bond), halo "
            "planes exchanged through HBM as bits")
    if engine in ("pslab", "slab", "lattice"):
        fmt = {"pslab": "slabs resident in LDS as int8 spins, one byte of bond signs per site",
               "slab": "int8 spins [replica][site], slabs staged in LDS each sweep, one byte of "
                       "bond signs per site",
               "lattice": "int8 spins [position][replica], implicit periodic neighbours, one byte "
                          "of bond signs per site"}[engine]
        return (d + 2) + 1.0, fmt
    deg = np.diff(np.asarray(enc.row_ptr))
    w_bytes = 4.0 if enc.weight.dtype.itemsize == 1 and model.n < (1 << 24) else 8.0
    b_h = 0.0 if enc.h is None else float(enc.h.dtype.itemsize)
    if engine == "lds":
        # the first G edge records of every row (G = 2, 3 or 4: the smallest covering three rows
        # in four), spins in LDS
        chunk = next((g for g in (2, 3, 4) if np.mean(deg <= g) >= 0.75), 4)
        b_s, fmt = 1.0, "int8 spins [position][replica] in LDS"
    elif engine == "global":
        chunk = 4 if np.mean(deg <= 4) >= 0.75 else 8
        if R <= 32:
            b_s, fmt = 0.125, "spins as one 32-bit word per position (one bit per replica)"
        else:
            b_s, fmt = 1.0, "int8 spins [position][replica] in HBM"
    else:   # coop: rows and edge records of each workgroup's slice in its LDS
        chunk = 0
        b_s, fmt = 1.0, "int8 spins [position][replica] in HBM, each workgroup's rows in its LDS"
    b_ptr = 8.0 if deg.max(initial=0) > chunk else 0.0
    if chunk:
        fmt += (f"; the first {chunk} edge records of each row ({int(w_bytes)} bytes: column and "
                f"{enc.weight.dtype.name} weight) in a zero-padded row-major copy read with 16-byte "
                f"loads" + ("; longer rows finished from packed CSR records" if b_ptr else ""))
    if enc.weight.dtype.kind == "f":
        fmt += "; float fields summed in double"
    return b_s * (d + 1) + b_s + d * w_bytes + b_ptr + b_h, fmt


def lds_bytes_per_attempt(kernel: str, model) -> float | None:
    """LDS bytes one attempt reads and writes in the kernels whose state lives in LDS."""
    d = model.to_csr()[1].size / model.n
    if kernel == "pslab_bits_run":
        # per 32-site row word (16 sites of the colour): the word, its two row neighbours, 2 (d-2)
        # other neighbour words, d bond words, four 16-byte uniform blocks, one store
        words = 3 + (d - 2) + d
        return (4.0 * words + 64.0 + 4.0) / 16.0
    if kernel == "pt_lds_kernel":
        return 1.0 + d + 0.5   # own spin, d neighbour spins, a flip store at most every other attempt
    return None


def median(v):
    return float(statistics.median(v))


_probe_cache = {}


def lds_probe() -> dict:
    """This device's LDS read rate (16-byte and 1-byte reads), measured once per process."""
    if _probe_cache:
        return _probe_cache
    lib = ctypes.CDLL(os.path.join(ROOT, "drift_amd", "libdriftprobe.so"))
    lib.drift_probe_lds.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    lib.drift_probe_lds.restype = ctypes.c_int
    out = {}
    for kind, name in ((0, "b128"), (1, "u8")):
        sec, byt, ins = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        rc = lib.drift_probe_lds(kind, 4096, ctypes.byref(sec), ctypes.byref(byt), ctypes.byref(ins))
        if rc != 0:
            raise RuntimeError(f"LDS probe failed ({rc})")
        out[name] = {"GBs": byt.value / sec.value / 1e9,
                     "wave_reads_per_s": ins.value / sec.value, "seconds": sec.value}
    _probe_cache.update(out)
    return out


def limiter(kernel: str):
    """What limits a kernel, from its rocprofv3 counters (profiles/limiters.json), or None."""
    path = os.path.join(PROFILES, "limiters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(kernel)


def traffic_per_attempt(kernel: str):
    """Counted HBM bytes per attempt of a kernel from its rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (profiles/limiters.json), with the profile it comes from."""
    lim = limiter(kernel)
    if not lim or "hbm_bytes_per_attempt" not in lim:
        return None, None
    return lim["hbm_bytes_per_attempt"], lim.get("source")


CUS, CLOCK_HZ = 256, 2.4e9   # MI355X compute units and peak engine clock


def roofline(kernel, B, B_survey, attempts_per_launch, launch_s, model=None):
    """The contract's HBM roofline at the declared storage format, with the ceilings that can bind
    a kernel whose state is on chip beside it: HBM traffic as counted by rocprofv3, VALU issue
    (one wave instruction per CU per cycle) and LDS issue (the probe's rate of wave-level LDS
    reads), each from the kernel's counter-measured instructions per attempt times the rate
    measured here.  `binding` names the ceiling the kernel sits closest to."""
    rate = attempts_per_launch / launch_s
    achieved = B * rate / 1e9
    line = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kernel, "bytes_per_attempt": round(B, 4),
            "attempts_per_launch": attempts_per_launch, "avg_launch_us": round(launch_s * 1e6, 3)}
    if B_survey is not None:
        line["bytes_per_attempt_plan"] = B_survey
        line["frac_plan"] = round(B_survey * rate / 1e9 / HBM_PEAK_GBS, 4)
    ceilings = {"hbm_algorithmic": line["frac"]}
    tpa, src = traffic_per_attempt(kernel)
    if tpa is not None:
        line["traffic"] = round(tpa * attempts_per_launch)
        line["traffic_source"] = src
        ceilings["hbm_counted"] = round(tpa * rate / (HBM_PEAK_GBS * 1e9), 4)
    lim = limiter(kernel)
    if lim:
        line["limiter"] = {k: v for k, v in lim.items() if k != "hbm_bytes_per_attempt"}
        if "valu_insts_per_attempt" in lim:
            ceilings["valu_issue"] = round(lim["valu_insts_per_attempt"] * rate / 64 / (CUS * CLOCK_HZ), 4)
        if "lds_insts_per_attempt" in lim and lds_bytes_per_attempt(kernel, model) is not None:
            ceilings["lds_issue"] = round(lim["lds_insts_per_attempt"] * rate / 64
                                          / lds_probe()["u8"]["wave_reads_per_s"], 4)
    lb = lds_bytes_per_attempt(kernel, model) if model is not None else None
    if lb is not None:
        peak = lds_probe()["b128"]["GBs"]
        ach = lb * rate / 1e9
        line["lds"] = {"bytes_per_attempt": round(lb, 3), "achieved_GBs": round(ach, 1),
                       "peak_GBs_measured": round(peak, 1), "frac": round(ach / peak, 4)}
    line["ceilings"] = ceilings
    measured = {k: v for k, v in ceilings.items() if k != "hbm_algorithmic"}
    top = max(measured, key=measured.get) if measured else None
    if top is not None and measured[top] >= 0.6:
        line["binding"] = top
    elif lim:
        line["binding"] = "latency (no throughput ceiling above 0.6; see limiter)"
    if ceilings["hbm_algorithmic"] > 1.0:
        line["note"] = ("frac above 1: the declared format counts bytes that stay on chip (spins in "
                        "LDS, edge records re-read from L2 every sweep), so HBM cannot be the "
                        "bound; the counted traffic and the issue ceilings show what is")
    return line


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


def cpu_baseline(model, enc, R, beta, S, budget_s: float):
    """The oracle's chain (its -O3 build, OpenMP over replicas) on this host's cores over a bounded
    number of rounds of the same workload, and on one thread over a fifth of the budget."""
    import oracle

    cores = len(os.sched_getaffinity(0))
    k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
    kw = dict(seed=1, R=R, sweeps_per_round=S, beta=beta, fast=True)
    probe = oracle.pt_run_encoded(enc, (cptr, cs), n_rounds=2, n_threads=cores, **kw)
    per_round = max(probe["seconds"] / 2, 1e-6)
    rounds = int(max(2, min(20000, budget_s / per_round)))
    run = oracle.pt_run_encoded(enc, (cptr, cs), n_rounds=rounds, n_threads=cores, **kw)
    attempts = R * model.n * S * rounds
    rounds1 = int(max(1, min(rounds, budget_s / 5 / (per_round * min(cores, R)))))
    one = oracle.pt_run_encoded(enc, (cptr, cs), n_rounds=rounds1, n_threads=1, **kw)
    return {
        "value": attempts / run["seconds"] / 1e9, "unit": "attempts/ns", "cores": min(cores, R),
        "kind": "port",
        "sample": f"same workload, {rounds} rounds of all {R} replicas ({attempts:.3g} attempts, "
                  f"{run['seconds']:.1f} s, OpenMP over replicas, -O3 build of the oracle)",
        "value_1_thread": R * model.n * S * rounds1 / one["seconds"] / 1e9,
        "sample_1_thread": f"{rounds1} rounds on one thread ({one['seconds']:.1f} s)",
        "cpu_model": cpu_model(), "host_cores": cores,
    }


def cpu_baseline_batch(models, R, beta, budget_s: float):
    """The oracle's chain on a sample of the batch, instances one after another (OpenMP over the
    replicas of each), sized to about budget_s seconds."""
    import oracle
    from drift_amd import gpu

    cores = len(os.sched_getaffinity(0))
    tot_s, tot_a, used = 0.0, 0, 0
    rounds = 50
    for m in models:
        enc = gpu.encode_problem(m)
        k, cptr, cs = oracle.greedy_coloring(enc.row_ptr, enc.col_idx)
        o = oracle.pt_run_encoded(enc, (cptr, cs), seed=1, R=R, n_rounds=rounds, sweeps_per_round=1,
                                  beta=beta, n_threads=cores, fast=True)
        tot_s += o["seconds"]
        tot_a += R * m.n * rounds
        used += 1
        if tot_s >= budget_s:
            break
    return {"value": tot_a / tot_s / 1e9, "unit": "attempts/ns", "cores": min(cores, R),
            "kind": "port",
            "sample": f"first {used} instances of the batch, {rounds} sweeps each, one after another "
                      f"(OpenMP over replicas, -O3 build of the oracle; {tot_a:.3g} attempts, "
                      f"{tot_s:.1f} s)",
            "cpu_model": cpu_model(), "host_cores": cores}


def bench_c4(steps, warmup, repeats, cpu_budget, count=4096):
    from drift_amd import gpu
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    log(f"c4: building {count} instances")
    models = c4_instances(count)
    R = 16
    kw = dict(n_replicas=R, T_min=0.2, T_max=5.0, sweeps_per_round=1)
    log(f"c4: warm-up {warmup} steps")
    gpu.parallel_tempering_gpu_batch(models, n_rounds=max(warmup, 1), **kw)
    total_n = sum(m.n for m in models)
    attempts = R * total_n * steps
    secs, walls = [], []
    for r in range(repeats):
        t0 = time.perf_counter()
        res = gpu.parallel_tempering_gpu_batch(models, n_rounds=steps, **kw)
        walls.append(time.perf_counter() - t0)
        secs.append(res[0].seconds)
        log(f"c4: timed run {r + 1}/{repeats}: {attempts / secs[-1] / 1e9:.1f} attempts/ns")
    med = median(secs)
    enc0 = gpu.encode_problem(models[0])
    # every instance is 3-regular with a field: the per-attempt bytes of one apply to all
    B, fmt = declared_bytes(models[0], enc0, "lds", R)
    out = {
        "value": round(attempts / med / 1e9, 3), "steps": steps, "warmup": warmup,
        "ms_per_step": round(med * 1e3 / steps, 5), "dtype": "f32",
        "config": {"workload": C4_DESC, "instances": count, "total_spins": int(total_n),
                   "attempts_per_step": R * total_n, "format": fmt, "engine": res[0].engine,
                   "wall_s_median": round(median(walls), 3)},
        "timing": {"repeats": repeats, "runs_ms": [round(x * 1e3, 4) for x in secs],
                   "median_ms": round(med * 1e3, 4), "window": "the launch that runs the rounds "
                   "(initial states drawn by a launch of their own, untimed)"},
        "roofline": roofline("pt_lds_kernel", B, SURVEY_BYTES["c4"], attempts, med, models[0]),
        "cpu_baseline": None,
    }
    if cpu_budget > 0:
        log("c4: CPU baseline")
        out["cpu_baseline"] = cpu_baseline_batch(models, R, ladder_betas(geometric_ladder(R, 0.2, 5.0)),
                                                 cpu_budget)
    return out


def bench_single(name, steps, warmup, repeats, cpu_budget, plain, engine="auto", block=0,
                 autotune=False, kernel_timing_steps=200):
    from drift_amd import gpu

    model, R, beta, S, desc = workload(name)
    enc = gpu.encode_problem(model)
    kw = dict(n_replicas=R, betas=beta, sweeps_per_round=S, seed=1, engine=engine,
              block_threads=block, record_history=False, plain_launch=plain,
              autotune=autotune and engine == "auto" and block == 0)
    log(f"{name}: warm-up {warmup} steps")
    gpu.parallel_tempering_gpu(model, n_rounds=max(warmup, 1), **kw)
    attempts = R * model.n * S * steps
    secs, walls = [], []
    for r in range(repeats):
        t0 = time.perf_counter()
        res = gpu.parallel_tempering_gpu(model, n_rounds=steps, **kw)
        walls.append(time.perf_counter() - t0)
        secs.append(res.seconds)
        log(f"{name}: timed run {r + 1}/{repeats}: {attempts / secs[-1] / 1e9:.2f} attempts/ns "
            f"({res.engine})")
    med = median(secs)
    kern = kernel_name(res.engine, model, R)
    B, fmt = declared_bytes(model, enc, res.engine, R)
    if res.engine in ("global", "lattice", "slab"):
        # several launches per step: the dominant kernel's average launch in the timed region is
        # the region's time per step times that kernel's share of device time (from a run that
        # brackets every launch with events) over its launches per step
        kt = gpu.parallel_tempering_gpu(model, n_rounds=kernel_timing_steps, time_kernels=True, **kw)
        per_step = kt.kernel_launches / kernel_timing_steps
        share = kt.kernel_seconds / (kt.kernel_seconds + (kt.other_kernel_seconds or 0.0))
        launch_s = med / steps * share / per_step
        launch_attempts = kt.kernel_attempts / kt.kernel_launches
    else:
        launch_s, launch_attempts = med, attempts   # one launch runs the whole timed region
    out = {
        "value": round(attempts / med / 1e9, 3), "steps": steps, "warmup": warmup,
        "ms_per_step": round(med * 1e3 / steps, 5),
        "dtype": "int32" if enc.is_integer else "f32",
        "config": {"workload": desc, "n": model.n, "replicas": R, "sweeps_per_step": S,
                   "attempts_per_step": R * model.n * S, "format": fmt, "engine": res.engine,
                   "wall_s_median": round(median(walls), 3)},
        "timing": {"repeats": repeats, "runs_ms": [round(x * 1e3, 4) for x in secs],
                   "median_ms": round(med * 1e3, 4)},
        "roofline": roofline(kern, B, SURVEY_BYTES.get(name), launch_attempts, launch_s, model),
        "cpu_baseline": None,
    }
    if cpu_budget > 0:
        log(f"{name}: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(model, enc, R, beta, S, cpu_budget)
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def headline(sub: dict) -> dict:
    return {
        "metric": "spin-flips/ns on 1 GPU", "value": sub["value"], "unit": "attempts/ns",
        "n_gpus": 1, "steps": sub["steps"], "warmup": sub["warmup"],
        "ms_per_step": sub["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": sub["dtype"], "data": "synthetic", "config": sub["config"],
        "roofline": sub["roofline"], "cpu_baseline": sub["cpu_baseline"], "timing": sub["timing"],
    }


SUB_DEFAULT_STEPS = {"c2": 10000, "c3": 1000, "c1": 20000, "published": 400}


def bench_throughput(args) -> int:
    cpu = 0.0 if args.no_cpu_baseline else args.cpu_budget
    if args.workload == "c4":
        print(json.dumps(headline(bench_c4(args.steps, args.warmup, args.repeats, cpu,
                                           args.instances))), flush=True)
        return 0
    if args.workload in ("c2", "c3", "c1", "published"):
        sub = bench_single(args.workload, args.steps, args.warmup, args.repeats, cpu,
                           args.plain_launch, args.engine, args.block, args.autotune)
        print(json.dumps(headline(sub)), flush=True)
        return 0
    # default: C4 headline with the other configs as sub-results
    line = headline(bench_c4(args.steps, args.warmup, args.repeats, cpu, args.instances))
    subs = {}
    for name in ("c2", "c3", "published"):
        steps = args.sub_steps or SUB_DEFAULT_STEPS[name]
        sub = bench_single(name, steps, min(args.warmup, steps) or 1, args.repeats,
                           cpu, args.plain_launch)
        if name == "published":
            sub["context"] = {"reference_rtx5060ti_attempts_per_ns": PUBLISHED_RTX5060TI,
                              "source": "docs/results/PHASE14-results.md:140-142",
                              "note": "a different card; context, not a target"}
        subs[name] = sub
    line["sub_results"] = subs
    line["lds_probe"] = {k: {"GBs": round(v["GBs"], 1),
                             "wave_reads_per_s": float(f"{v['wave_reads_per_s']:.4g}")}
                         for k, v in lds_probe().items()}
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=5, help="timed runs; the value is their median")
    ap.add_argument("--workload", default="all",
                    choices=("all", "c4", "c2", "c3", "c1", "published", "ladder", "ladder-c5"))
    ap.add_argument("--sub-steps", type=int, default=0,
                    help="steps of the sub-result workloads (0: their defaults)")
    ap.add_argument("--engine", default="auto")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--autotune", action="store_true",
                    help="with --engine auto: time every candidate engine first")
    ap.add_argument("--plain-launch", action="store_true",
                    help="ordinary dispatch of the persistent engines (for profiler commands)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
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
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this engine has no "
              "multi-GPU path; the sharding design is described in DESIGN.md", file=sys.stderr)
        return 2

    from drift_amd import gpu

    if not gpu.gpu_available():
        print("bench.py: GPU engine not available (no gfx950 device or library not built)",
              file=sys.stderr)
        return 1
    if args.workload == "ladder":
        return bench_ladder(args)
    if args.workload == "ladder-c5":
        return bench_ladder_c5(args)
    return bench_throughput(args)


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


if __name__ == "__main__":
    raise SystemExit(main())
