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
  c5_batch   64 random 3-regular float instances of N = 65536, R = 32: the split engine's batch
  single3    one random 3-regular float instance of N = 65536, R = 32, T 0.1 .. 3.0: the default
             engine for it, one-slot split groups in 8 parts
  time_to_target  the metric's second part on a sample of the learned-ladder evaluation's held-out
             set (8 instances per family, N ~ 65536, cap 10^5 sweeps): wall time of one call per
             instance on the geometric ladder (each on its autotuned engine, tuned untimed first)
             over one batched call on the learned ladders, and how many reach their best-known
             energies (--no-time-to-target leaves it out)

Rooflines count the bytes one attempt touches at the storage format the engine declares (no
credit for on-chip reuse); the fixed per-attempt figure of the measurement plan (8.75 B for C2,
50 B for C3, 45 B for C4) is shown beside it.  Kernels whose state lives in LDS are also put
against this device's LDS read rate, measured by a probe kernel at the start of the run.  What
limits each kernel is taken from its rocprofv3 counters (profiles/limiters.json).

    python bench.py [--steps K] [--warmup W] [--repeats 5] [--workload all|c4|c2|c3|c1|published]
                    [--no-cpu-baseline] [--plain-launch] [--no-autotune]

Before the timed runs, the library's autotuners time the candidate engines and launch shapes of
each workload (the batch: its launch shapes on the batch itself) and keep the default unless a
candidate is more than 3% faster; --no-autotune skips that.  --plain-launch starts the persistent
engines (pslab, coop, split) with an ordinary dispatch, for profiler commands; timed runs use the
cooperative launch.  --workload ladder-c5 measures time to target
with the learned ladder instead, in three stages (--stage train, bestknown, eval).
--workload single-heldout times one held-out instance per family (N ~ 65536) on each engine that
can run it alone, the split engine with its groups in parts among them; --workload parts-sweep
does the same across sizes and colour-class counts (the evidence for the default engine's rule).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
PROFILES = os.path.join(ROOT, "profiles")
PUBLISHED_RTX5060TI = 1.29   # attempts/ns, docs/results/PHASE14-results.md:140-142

# the fixed per-attempt byte figures of the measurement plan, per config
PLAN_BYTES = {"c2": 8.75, "c3": 50.0, "c4": 45.0}


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
    if name == "single3":
        from drift_amd import ladder

        T_min, T_max = ladder.MIX_FAMILIES["rrg"]
        return (ladder.mix_instance("rrg", 901_000, 65536), 32,
                ladder_betas(geometric_ladder(32, T_min, T_max)), 1,
                "one random 3-regular instance, N=65536, J ~ N(0,1) and h ~ N(0,0.1^2) in "
                "float32 (the first held-out 3-regular instance of the learned-ladder mix), R=32, "
                "T 0.1..3.0 geometric, one sweep and one exchange per step")
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


def kernel_name(engine: str, model, R: int, parts: int = 0) -> str:
    if engine == "split" and parts > 1:
        return "pt_split_kernel_parts"
    return {"global": "g_color_bits" if R <= 32 else "g_color_step",
            "lattice": "g_lattice_step", "slab": "slab_sweep", "coop": "coop_run",
            # bit slabs (side a multiple of 32) run without a grid barrier when no per-round history
            # or trace is recorded, as here
            "pslab": "pslab_bits_nbr" if lattice_side(model.n) % 32 == 0 else "pslab_run",
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
    if engine in ("lds", "split"):
        # the first G edge records of every row (G = 2, 3, 4, or 6 for int8 couplings: the
        # smallest covering three rows in four, else the widest), spins in LDS
        widths = (2, 3, 4, 6) if w_bytes == 4.0 else (2, 3, 4)
        chunk = next((g for g in widths if np.mean(deg <= g) >= 0.75), widths[-1])
        b_s, fmt = 1.0, ("int8 spins [position][replica] in LDS" if engine == "lds" else
                         "int8 spins [position][slot] of each workgroup's group of replica slots "
                         "in its LDS; slot energies traded through round-tagged mailboxes in HBM "
                         "after every round")
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
    if kernel in ("pslab_bits_run", "pslab_bits_nbr"):
        # per 32-site row word (16 sites of the colour): the word, its two row neighbours, 2 (d-2)
        # other neighbour words, d bond words, four 16-byte uniform blocks, one store
        words = 3 + (d - 2) + d
        return (4.0 * words + 64.0 + 4.0) / 16.0
    if kernel in ("pt_lds_kernel", "pt_split_kernel"):
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


def roofline(kernel, B, B_plan, attempts_per_launch, launch_s, model=None):
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
    if B_plan is not None:
        line["bytes_per_attempt_plan"] = B_plan
        line["frac_plan"] = round(B_plan * rate / 1e9 / HBM_PEAK_GBS, 4)
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


def bench_c4(steps, warmup, repeats, cpu_budget, count=4096, autotune=False):
    from drift_amd import gpu
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    log(f"c4: building {count} instances")
    models = c4_instances(count)
    R = 16
    kw = dict(n_replicas=R, T_min=0.2, T_max=5.0, sweeps_per_round=1, autotune=autotune)
    log(f"c4: warm-up {warmup} steps" + (" (the batch autotuner times launch shapes first)"
                                          if autotune else ""))
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
                   "block_threads": res[0].block_threads, "autotune": bool(autotune),
                   "wall_s_median": round(median(walls), 3)},
        "timing": {"repeats": repeats, "runs_ms": [round(x * 1e3, 4) for x in secs],
                   "median_ms": round(med * 1e3, 4), "window": "the launch that runs the rounds "
                   "(initial states drawn by a launch of their own, untimed)"},
        "roofline": roofline("pt_lds_kernel", B, PLAN_BYTES["c4"], attempts, med, models[0]),
        "cpu_baseline": None,
    }
    if cpu_budget > 0:
        log("c4: CPU baseline")
        out["cpu_baseline"] = cpu_baseline_batch(models, R, ladder_betas(geometric_ladder(R, 0.2, 5.0)),
                                                 cpu_budget)
    return out


C5_BATCH_DESC = ("64 random 3-regular instances of N = 65536 (the held-out size of the learned-ladder "
                 "evaluation), J ~ N(0,1) and h ~ N(0,0.1^2) in float32, R=32 each, beta 0.2..5.0 "
                 "geometric; too large for one workgroup, so the batch runs on the split engine in "
                 "one persistent launch; one sweep and one exchange of every instance per step")
C5_BATCH_ER_DESC = ("the learned-ladder evaluation's held-out sparse Erdos-Renyi instances of N=65536 "
                    "(mean degree 3..12, weights in {+-1,+-2,+-3}), R=32, beta 0.2..5.0 geometric, "
                    "all in one persistent split-engine launch; one sweep and one exchange of every "
                    "instance per step")


def bench_c5_batch(steps, warmup, repeats, cpu_budget, count=64, autotune=False, plain=False,
                   family="rrg"):
    """Batched instances too large for one workgroup's LDS (the split engine): random 3-regular
    float graphs, or (family "er") the learned-ladder evaluation's held-out sparse Erdos-Renyi
    graphs (int8 weights, mean degree 3 to 12)."""
    from drift_amd import gpu, ladder
    from drift_amd.builders.lattice import random_regular_gaussian
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    log(f"c5-batch: building {count} {family} instances of N = 65536")
    if family == "er":
        models = [ladder.mix_instance("er", 902_000 + k, 65536) for k in range(count)]
    else:
        models = [random_regular_gaussian(65536, seed=60_000 + i) for i in range(count)]
    R = 32
    kw = dict(n_replicas=R, T_min=0.2, T_max=5.0, sweeps_per_round=1, autotune=autotune,
              plain_launch=plain)
    log(f"c5-batch: warm-up {warmup} steps")
    gpu.parallel_tempering_gpu_batch(models, n_rounds=max(warmup, 1), **kw)
    total_n = sum(m.n for m in models)
    attempts = R * total_n * steps
    secs, walls = [], []
    for r in range(repeats):
        t0 = time.perf_counter()
        res = gpu.parallel_tempering_gpu_batch(models, n_rounds=steps, **kw)
        walls.append(time.perf_counter() - t0)
        secs.append(res[0].seconds)
        log(f"c5-batch: timed run {r + 1}/{repeats}: {attempts / secs[-1] / 1e9:.1f} attempts/ns "
            f"({res[0].engine}, {res[0].slots_per_group} slots per workgroup)")
    med = median(secs)
    enc0 = gpu.encode_problem(models[0])
    B, fmt = declared_bytes(models[0], enc0, res[0].engine, R)
    kern = (("pt_split_kernel_er_batch" if family == "er" else "pt_split_kernel")
            if res[0].engine == "split" else kernel_name(res[0].engine, models[0], R))
    out = {
        "value": round(attempts / med / 1e9, 3), "steps": steps, "warmup": warmup,
        "ms_per_step": round(med * 1e3 / steps, 5), "dtype": "f32",
        "config": {"workload": C5_BATCH_DESC if family != "er" else C5_BATCH_ER_DESC,
                   "instances": count, "total_spins": int(total_n),
                   "attempts_per_step": R * total_n, "format": fmt, "engine": res[0].engine,
                   "slots_per_group": res[0].slots_per_group, "block_threads": res[0].block_threads,
                   "autotune": bool(autotune), "wall_s_median": round(median(walls), 3)},
        "timing": {"repeats": repeats, "runs_ms": [round(x * 1e3, 4) for x in secs],
                   "median_ms": round(med * 1e3, 4), "window": "the launch that runs the rounds "
                   "(initial states drawn by a launch of their own, untimed)"},
        "roofline": roofline(kern, B, PLAN_BYTES["c4"], attempts, med, models[0]),
        "cpu_baseline": None,
    }
    if cpu_budget > 0:
        log("c5-batch: CPU baseline")
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
    kern = kernel_name(res.engine, model, R, res.parts_per_group or 0)
    B, fmt = declared_bytes(model, enc, res.engine, R)
    if res.engine == "split" and (res.parts_per_group or 0) > 1:
        fmt += (f"; each group's positions split over {res.parts_per_group} workgroups, which trade "
                "every colour class's new spins as step-tagged words of 32 spin bits in HBM")
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
                   "block_threads": res.block_threads, "autotune": bool(kw["autotune"]),
                   "slots_per_group": res.slots_per_group, "parts_per_group": res.parts_per_group,
                   "wall_s_median": round(median(walls), 3)},
        "timing": {"repeats": repeats, "runs_ms": [round(x * 1e3, 4) for x in secs],
                   "median_ms": round(med * 1e3, 4)},
        "roofline": roofline(kern, B, PLAN_BYTES.get(name), launch_attempts, launch_s, model),
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


SUB_DEFAULT_STEPS = {"c2": 10000, "c3": 1000, "c1": 20000, "published": 400, "c5_batch": 200,
                     "single3": 500}


def bench_throughput(args) -> int:
    cpu = 0.0 if args.no_cpu_baseline else args.cpu_budget
    if args.workload == "c4":
        print(json.dumps(headline(bench_c4(args.steps, args.warmup, args.repeats, cpu,
                                           args.instances, args.autotune))), flush=True)
        return 0
    if args.workload == "c5-batch":
        print(json.dumps(headline(bench_c5_batch(args.steps, args.warmup, args.repeats, cpu,
                                                 count=args.batch_count,
                                                 autotune=args.autotune,
                                                 plain=args.plain_launch,
                                                 family=args.batch_family))), flush=True)
        return 0
    if args.workload in ("c2", "c3", "c1", "published", "single3"):
        sub = bench_single(args.workload, args.steps, args.warmup, args.repeats, cpu,
                           args.plain_launch, args.engine, args.block, args.autotune)
        print(json.dumps(headline(sub)), flush=True)
        return 0
    # default: C4 headline with the other configs as sub-results
    line = headline(bench_c4(args.steps, args.warmup, args.repeats, cpu, args.instances,
                             args.autotune))
    subs = {}
    failed = []

    def sub_result(name, run):
        # a sub-result that raises is recorded and ends the GPU work; the headline still prints
        if failed:
            subs[name] = {"skipped": f"not run after the failure of {failed[0]}"}
            return
        try:
            subs[name] = run()
        except Exception as e:
            failed.append(name)
            subs[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"sub-result {name} failed: {type(e).__name__}: {e}")

    def single(name):
        steps = args.sub_steps or SUB_DEFAULT_STEPS[name]
        sub = bench_single(name, steps, min(args.warmup, steps) or 1, args.repeats,
                           cpu, args.plain_launch, autotune=args.autotune)
        if name == "published":
            sub["context"] = {"reference_rtx5060ti_attempts_per_ns": PUBLISHED_RTX5060TI,
                              "source": "docs/results/PHASE14-results.md:140-142",
                              "note": "a different card; context, not a target"}
        return sub

    for name in ("c2", "c3", "published", "single3"):
        sub_result(name, lambda name=name: single(name))
    sub_result("c5_batch", lambda: bench_c5_batch(args.sub_steps or SUB_DEFAULT_STEPS["c5_batch"],
                                                  min(args.warmup, 20) or 1, args.repeats, cpu,
                                                  autotune=args.autotune,
                                                  plain=args.plain_launch))
    if not args.no_time_to_target:
        sub_result("time_to_target", bench_ttt_sample)
    line["sub_results"] = subs
    if not failed:
        line["lds_probe"] = {k: {"GBs": round(v["GBs"], 1),
                                 "wave_reads_per_s": float(f"{v['wave_reads_per_s']:.4g}")}
                             for k, v in lds_probe().items()}
    print(json.dumps(line), flush=True)
    return 0


def _write_json(path, obj):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(obj, f, indent=1)


def c5_train(args) -> dict:
    """Labels, features and the network.  Each training instance runs every ladder shape of
    ladder.BUDGET_GAMMAS for --label-budget sweeps on the same streams; its label is the shape with
    the lowest best energy (the earliest on ties).  The 16 features come from the instance and a
    200-sweep run; the network is trained by the library's HIP kernels and saved to --net."""
    from drift_amd import ladder

    t_all = time.perf_counter()
    items = ladder.training_instances(args.train_per_family)
    feats, labels, picks, sizes = [], [], {}, []
    for fam, (T_min, T_max) in ladder.MIX_FAMILIES.items():
        models = [m for f, m in items if f == fam]
        t0 = time.perf_counter()
        for c0 in range(0, len(models), 256):
            chunk = models[c0:c0 + 256]
            x, pick = ladder.budget_labels(chunk, T_min=T_min, T_max=T_max, R=32,
                                           budget=args.label_budget, seeds=range(c0, c0 + len(chunk)))
            labels.append(x)
            picks.setdefault(fam, []).extend(int(k) for k in pick)
            feats.append(ladder.instance_features(chunk, T_min=T_min, T_max=T_max, R=32))
            log(f"c5 train: {fam} labels {c0 + len(chunk)}/{len(models)} "
                f"({time.perf_counter() - t0:.1f} s)")
        sizes += [m.n for m in models]
    X, Y = np.concatenate(feats), np.concatenate(labels)
    net = ladder.LadderNet(seed=11)
    t0 = time.perf_counter()
    losses = net.fit(X, Y, epochs=args.epochs, lr=3e-3)
    net.save(args.net)
    os.makedirs(args.out, exist_ok=True)
    np.savez(os.path.join(args.out, "c5_train_data.npz"), features=X, labels=Y)
    return {"stage": "train", "instances": int(X.shape[0]), "n_min": int(min(sizes)),
            "n_max": int(max(sizes)), "label_budget_sweeps": args.label_budget,
            "label_shapes": [float(g) for g in ladder.BUDGET_GAMMAS],
            "chosen_shape": {f: {str(g): p.count(k) for k, g in enumerate(ladder.BUDGET_GAMMAS)}
                             for f, p in picks.items()},
            "train_loss": [float(losses[0]), float(losses[-1])], "epochs": args.epochs,
            "train_s": round(time.perf_counter() - t0, 1), "net": os.path.basename(args.net),
            "total_s": round(time.perf_counter() - t_all, 1)}


# best-known runs: four independent chains per held-out instance, two on the geometric ladder and
# two on the power ladder of exponent 0.7 (a shape neither favours), each --cap x 10 sweeps
BESTKNOWN_RUNS = ((101, 1.0), (102, 1.0), (103, 0.7), (104, 0.7))


def c5_bestknown(args) -> dict:
    """Best-known energies of the held-out instances, from independent runs ten times the cap."""
    from drift_amd import ladder

    t_all = time.perf_counter()
    held = ladder.heldout_instances(args.test_per_family, args.test_n)
    rows = {}
    states = {}
    rounds = 10 * args.cap
    fix_path = os.path.join(args.out, "c5_heldout_bestknown.json")
    st_path = os.path.join(args.out, "c5_heldout_states.npz")
    families = args.families.split(",") if args.families else list(ladder.MIX_FAMILIES)
    if args.families and os.path.exists(fix_path):
        # a run over some families adds to the rows of the same held-out set recorded before
        with open(fix_path) as f:
            old = json.load(f)
        if (old["per_family"], old["test_n"], old["cap_sweeps"]) == (args.test_per_family, args.test_n, args.cap):
            rows = {r["index"]: r for r in old["rows"] if r["family"] not in families}
            with np.load(st_path) as z:
                states = {k: z[k] for k in z.files if int(k) in rows}
    for fam, (T_min, T_max) in ladder.MIX_FAMILIES.items():
        if fam not in families:
            continue
        idx = [i for i, (f, _) in enumerate(held) if f == fam]
        models = [held[i][1] for i in idx]
        for seed, gamma in BESTKNOWN_RUNS:
            b = ladder.betas_from_positions(ladder.power_positions(gamma), T_min, T_max)
            t0 = time.perf_counter()
            res = ladder.run_instances(models, n_replicas=32, T_min=T_min, T_max=T_max,
                                       n_rounds=rounds, seeds=[seed] * len(models),
                                       betas=[b] * len(models))
            for i, m, r in zip(idx, models, res):
                row = rows.setdefault(i, {"index": i, "family": fam, "n": m.n, "runs": []})
                row["runs"].append({"seed": seed, "ladder_power": gamma, "best_E": r.best_E,
                                    "best_round": r.best_round, "engine": r.engine})
                if "best_known_energy" not in row or r.best_E < row["best_known_energy"]:
                    row["best_known_energy"] = r.best_E
                    states[str(i)] = np.packbits(r.best_s < 0)
            log(f"c5 best-known: {fam} seed {seed} ({time.perf_counter() - t0:.1f} s)")
    out = [rows[i] for i in sorted(rows)]
    for r in out:
        r["runs_reaching_best"] = sum(x["best_E"] == r["best_known_energy"] for x in r["runs"])
    fix = {"description": "best-known energies of the held-out instances of the learned-ladder "
                          "evaluation: the lowest of four independent runs",
           "source": f"gpu-pt, {rounds} rounds of one sweep each, R=32, seeds and ladder shapes "
                     f"{list(BESTKNOWN_RUNS)}", "per_family": args.test_per_family,
           "test_n": args.test_n, "cap_sweeps": args.cap, "rows": out}
    os.makedirs(args.out, exist_ok=True)
    _write_json(fix_path, fix)
    np.savez_compressed(st_path, **states)
    return {"stage": "bestknown", "instances": len(out), "rounds": rounds,
            "total_s": round(time.perf_counter() - t_all, 1)}


def _reach(hist, target, rel):
    """First round (1-based) whose best energy is within rel of the target, or None."""
    tol = rel * max(1.0, abs(target)) + 1e-9 * max(1.0, abs(target))
    hit = np.nonzero(np.asarray(hist) <= target + tol)[0]
    return int(hit[0]) + 1 if hit.size else None


def c5_eval(args) -> dict:
    """Time to the best-known energy on the held-out instances, every run stopping at its target
    within --cap sweeps: (a) one call per instance on the geometric ladder, as the reference runs
    it, each on the engine and launch shape the autotuner picks for its shape class (the tuning
    runs untimed before the family's timed calls); (b) one batched call on the learned ladders;
    (c) one batched call on the geometric ladders (the batching alone).  Wall time is the library
    calls' host time; the instances' encoding is the same for all three and is not counted.  --families runs some families only and writes
    their rows to --out; --stage eval-merge then puts the parts together."""
    from drift_amd import gpu, ladder

    t_all = time.perf_counter()
    with open(args.bestknown) as f:
        bk = json.load(f)
    target = {r["index"]: r["best_known_energy"] for r in bk["rows"]}
    net = ladder.LadderNet(seed=11)
    net.load(args.net)
    cap = args.cap
    families = args.families.split(",") if args.families else list(ladder.MIX_FAMILIES)
    rows = []
    walls = {"geo_single": 0.0, "learned_batched": 0.0, "geo_batched": 0.0}
    for fi, (fam, (T_min, T_max)) in enumerate(ladder.MIX_FAMILIES.items()):
        if fam not in families:
            continue
        # the held-out instances of this family (ladder.heldout_instances, one family at a time)
        idx = [fi * bk["per_family"] + k for k in range(bk["per_family"])]
        models = [ladder.mix_instance(fam, 900_000 + 1000 * fi + k, bk["test_n"])
                  for k in range(bk["per_family"])]
        tg = [target[i] for i in idx]
        seeds = [4242 + i for i in idx]
        xs = net.predict_ladder(ladder.instance_features(models, T_min=T_min, T_max=T_max, R=32))
        learned = [ladder.betas_from_positions(x, T_min, T_max) for x in xs]
        kw = dict(n_replicas=32, T_min=T_min, T_max=T_max, n_rounds=cap, sweeps_per_round=1,
                  record_history=True)
        # (a) the reference's way: one instance per call, geometric ladder, on the tuned engine
        _tune_singles(models, seeds, tg, kw)
        log(f"c5 eval: {fam} engines tuned; single geometric")
        single = []
        for m, s, t in zip(models, seeds, tg):
            t0 = time.perf_counter()
            r = gpu.parallel_tempering_gpu(m, seed=s, target_energy=t, autotune=True, **kw)
            walls["geo_single"] += time.perf_counter() - t0
            single.append(r)
            if len(single) % 8 == 0:
                log(f"c5 eval: {fam} single geometric {len(single)}/{len(models)}")
        log(f"c5 eval: {fam} single geometric done; batched learned")
        t0 = time.perf_counter()
        lrn = gpu.parallel_tempering_gpu_batch(models, seeds=seeds, betas=learned, targets=tg, **kw)
        walls["learned_batched"] += time.perf_counter() - t0
        log(f"c5 eval: {fam} batched learned done; batched geometric")
        t0 = time.perf_counter()
        geo = gpu.parallel_tempering_gpu_batch(models, seeds=seeds, targets=tg, **kw)
        walls["geo_batched"] += time.perf_counter() - t0
        log(f"c5 eval: {fam} batched runs done")
        for i, m, t, a, b, c, x in zip(idx, models, tg, single, lrn, geo, xs):
            rows.append({"index": i, "family": fam, "n": m.n, "target": t,
                         "geo_single": {"rounds": a.rounds_run, "reached": a.best_E <= t + 1e-9 * abs(t),
                                        "best_E": a.best_E, "device_s": a.seconds,
                                        "engine": a.engine, "slots_per_group": a.slots_per_group,
                                        "parts_per_group": a.parts_per_group},
                         "learned_batched": {"rounds": b.rounds_run, "reached": b.best_E <= t + 1e-9 * abs(t),
                                             "best_E": b.best_E},
                         "geo_batched": {"rounds": c.rounds_run, "reached": c.best_E <= t + 1e-9 * abs(t),
                                         "best_E": c.best_E},
                         "within_1e-4": {k: _reach(r.best_history[:max(r.rounds_run, 1)], t, 1e-4)
                                         for k, r in (("geo", a), ("learned", b))},
                         "within_1e-3": {k: _reach(r.best_history[:max(r.rounds_run, 1)], t, 1e-3)
                                         for k, r in (("geo", a), ("learned", b))},
                         # a target both can reach: the energy the geometric run ends at
                         "to_geo_final": {k: _reach(r.best_history[:max(r.rounds_run, 1)], a.best_E, 0.0)
                                          for k, r in (("geo", a), ("learned", b))},
                         "learned_rung8_position": float(x[8]),
                         "learned_power": float(np.log(max(x[8], 1e-12)) / np.log(8 / 31))})

    part = {"cap": cap, "families": families, "walls": walls, "rows": rows,
            "net": os.path.basename(args.net), "total_s": round(time.perf_counter() - t_all, 1)}
    if args.families:
        os.makedirs(args.out, exist_ok=True)
        path = os.path.join(args.out, f"c5_eval_part_{'_'.join(families)}_cap{cap}.json")
        _write_json(path, part)
        log(f"c5 eval: rows of {families} written to {path}")
    return c5_eval_summary(part["rows"], part["walls"], cap, bk, part["net"], part["total_s"])


def c5_eval_batched(args) -> dict:
    """The evaluation's batched legs over every family at once: (b) one batched call on the
    learned ladders and (c) one on the geometric ladders for all held-out instances of the families
    in --parts (part files of --stage eval, whose one-call-per-instance leg (a) and its wall time
    are taken over as they are).  The library runs each call's instances in persistent launches
    by weight type, so one family's slow instances overlap another family's work instead of
    finishing alone.  Writes the complete rows to --out and returns the summary."""
    from drift_amd import gpu, ladder
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    t_all = time.perf_counter()
    with open(args.bestknown) as f:
        bk = json.load(f)
    single, walls_a, caps, nets = {}, 0.0, set(), set()
    for path in args.parts.split(","):
        with open(path) as f:
            part = json.load(f)
        caps.add(part["cap"])
        nets.add(part["net"])
        walls_a += part["walls"]["geo_single"]
        for r in part["rows"]:
            single[r["index"]] = r
    if len(caps) != 1 or len(nets) != 1:
        raise SystemExit("eval-batched: the parts were run with different caps or networks")
    cap = caps.pop()
    net = ladder.LadderNet(seed=11)
    net.load(args.net)
    models, idx, targets, learned, geometric, seeds, xs_all = [], [], [], [], [], [], []
    for fi, (fam, (T_min, T_max)) in enumerate(ladder.MIX_FAMILIES.items()):
        ks = [k for k in range(bk["per_family"]) if fi * bk["per_family"] + k in single]
        if not ks:
            continue
        fam_models = [ladder.mix_instance(fam, 900_000 + 1000 * fi + k, bk["test_n"]) for k in ks]
        xs = net.predict_ladder(ladder.instance_features(fam_models, T_min=T_min, T_max=T_max, R=32))
        geo_b = ladder_betas(geometric_ladder(32, T_min, T_max))
        for k, m, x in zip(ks, fam_models, xs):
            i = fi * bk["per_family"] + k
            models.append(m)
            idx.append(i)
            targets.append(single[i]["target"])
            learned.append(ladder.betas_from_positions(x, T_min, T_max))
            geometric.append(geo_b)
            seeds.append(4242 + i)
            xs_all.append(x)
    log(f"c5 eval-batched: {len(models)} instances in one batched call per ladder")
    kw = dict(n_replicas=32, n_rounds=cap, sweeps_per_round=1, record_history=True,
              slots_per_group=args.slots_per_group)
    t0 = time.perf_counter()
    lrn = gpu.parallel_tempering_gpu_batch(models, seeds=seeds, betas=learned, targets=targets, **kw)
    wall_b = time.perf_counter() - t0
    log(f"c5 eval-batched: learned done in {wall_b:.1f} s; geometric")
    if args.learned_only:   # leg (c) as the part files have it
        geo = [types.SimpleNamespace(rounds_run=single[i]["geo_batched"]["rounds"],
                                     best_E=single[i]["geo_batched"]["best_E"]) for i in idx]
        wall_c = sum(json.load(open(p))["walls"]["geo_batched"] for p in args.parts.split(","))
    else:
        t0 = time.perf_counter()
        geo = gpu.parallel_tempering_gpu_batch(models, seeds=seeds, betas=geometric,
                                               targets=targets, **kw)
        wall_c = time.perf_counter() - t0
        log(f"c5 eval-batched: geometric done in {wall_c:.1f} s")
    rows = []
    for i, t, b, c, x in zip(idx, targets, lrn, geo, xs_all):
        a = single[i]
        hb = b.best_history[:max(b.rounds_run, 1)]
        rows.append({"index": i, "family": a["family"], "n": a["n"], "target": t,
                     "geo_single": a["geo_single"],
                     "learned_batched": {"rounds": b.rounds_run, "reached": b.best_E <= t + 1e-9 * abs(t),
                                         "best_E": b.best_E},
                     "geo_batched": {"rounds": c.rounds_run, "reached": c.best_E <= t + 1e-9 * abs(t),
                                     "best_E": c.best_E},
                     "within_1e-4": {"geo": a["within_1e-4"]["geo"], "learned": _reach(hb, t, 1e-4)},
                     "within_1e-3": {"geo": a["within_1e-3"]["geo"], "learned": _reach(hb, t, 1e-3)},
                     "to_geo_final": {"geo": a["to_geo_final"]["geo"],
                                      "learned": _reach(hb, a["geo_single"]["best_E"], 0.0)},
                     "learned_rung8_position": float(x[8]),
                     "learned_power": float(np.log(max(x[8], 1e-12)) / np.log(8 / 31))})
    walls = {"geo_single": walls_a, "learned_batched": wall_b, "geo_batched": wall_c}
    part = {"cap": cap, "families": sorted({r["family"] for r in rows}), "walls": walls,
            "rows": rows, "net": nets.pop(), "total_s": round(time.perf_counter() - t_all, 1),
            "batched": "one call over every family"}
    os.makedirs(args.out, exist_ok=True)
    tag = "_".join(part["families"]) + (f"_rg{args.slots_per_group}" if args.slots_per_group else "")
    part["slots_per_group"] = args.slots_per_group
    path = os.path.join(args.out, f"c5_eval_part_batched_{tag}_cap{cap}.json")
    _write_json(path, part)
    log(f"c5 eval-batched: rows written to {path}")
    return c5_eval_summary(rows, walls, cap, bk, part["net"], part["total_s"])


def c5_eval_merge(args) -> dict:
    """The evaluation's summary over parts written by --stage eval --families (--parts)."""
    with open(args.bestknown) as f:
        bk = json.load(f)
    rows, walls, caps, nets, total = [], {}, set(), set(), 0.0
    for path in args.parts.split(","):
        with open(path) as f:
            part = json.load(f)
        caps.add(part["cap"])
        nets.add(part["net"])
        rows += part["rows"]
        total += part["total_s"]
        for k, v in part["walls"].items():
            walls[k] = walls.get(k, 0.0) + v
    if len(caps) != 1 or len(nets) != 1:
        raise SystemExit("eval-merge: the parts were run with different caps or networks")
    seen = [r["index"] for r in rows]
    if len(seen) != len(set(seen)):
        raise SystemExit("eval-merge: an instance appears in two parts")
    rows.sort(key=lambda r: r["index"])
    line = c5_eval_summary(rows, walls, caps.pop(), bk, nets.pop(), round(total, 1))
    line["parts"] = [os.path.basename(p) for p in args.parts.split(",")]
    return line


def c5_eval_summary(rows, walls, cap, bk, net_name, total_s) -> dict:
    from drift_amd import ladder

    def summary(sel):
        n = len(sel)
        out = {"instances": n}
        for k in ("geo_single", "learned_batched", "geo_batched"):
            out[f"reached_{k}"] = int(sum(r[k]["reached"] for r in sel))
            # a run that stopped at its target within half (a quarter of) the cap: what a run
            # capped there reaches
            out[f"reached_within_half_cap_{k}"] = int(sum(r[k]["reached"] and r[k]["rounds"] <= cap // 2
                                                          for r in sel))
            out[f"reached_within_quarter_cap_{k}"] = int(sum(r[k]["reached"] and r[k]["rounds"] <= cap // 4
                                                             for r in sel))
            out[f"median_rounds_{k}"] = float(np.median([r[k]["rounds"] for r in sel]))
        for tol in ("within_1e-4", "within_1e-3"):
            out[f"{tol}_geo"] = int(sum(r[tol]["geo"] is not None for r in sel))
            out[f"{tol}_learned"] = int(sum(r[tol]["learned"] is not None for r in sel))
        # sweeps to the geometric run's final energy (unreached: twice the cap)
        tg = np.array([r["to_geo_final"]["geo"] or 2 * cap for r in sel], dtype=float)
        tl = np.array([r["to_geo_final"]["learned"] or 2 * cap for r in sel], dtype=float)
        out["to_geo_final_reached_learned"] = int(sum(r["to_geo_final"]["learned"] is not None for r in sel))
        out["to_geo_final_median_sweeps_geo"] = float(np.median(tg))
        out["to_geo_final_median_sweeps_learned"] = float(np.median(tl))
        out["to_geo_final_sweeps_ratio_median"] = float(np.median(tg / tl))
        out["learned_lower_final_energy"] = int(sum(r["learned_batched"]["best_E"] < r["geo_single"]["best_E"] for r in sel))
        out["median_gap_to_best_known_geo"] = float(np.median([(r["geo_single"]["best_E"] - r["target"]) / abs(r["target"]) for r in sel]))
        out["median_gap_to_best_known_learned"] = float(np.median([(r["learned_batched"]["best_E"] - r["target"]) / abs(r["target"]) for r in sel]))
        return out

    total = summary(rows)
    speedup = walls["geo_single"] / max(walls["learned_batched"], 1e-9)
    return {
        "metric": "time-to-target energy, learned vs geometric ladder",
        "value": round(speedup, 3),
        "unit": "x less wall time to the best-known energies (batched + learned vs one call per "
                "instance + geometric, every run stopping at its target)",
        "n_gpus": 1, "steps": cap, "warmup": 0, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32/f32", "data": "synthetic",
        "config": {"workload": f"held-out instances of the learned-ladder mix at N ~ {bk['test_n']} "
                               "(3-D +/-J glass at the nearest cube, 3-regular N(0,1), sparse ER "
                               "mean degree 3..12), R=32",
                   "per_family": bk["per_family"], "cap_sweeps": cap,
                   "target": f"best-known energy: lowest of four independent runs of "
                             f"{10 * bk['cap_sweeps']} sweeps (" + bk["source"] + ")",
                   "net": net_name},
        "wall_s": {k: round(v, 3) for k, v in walls.items()},
        "speedup_ladder_only": round(walls["geo_batched"] / max(walls["learned_batched"], 1e-9), 3),
        "speedup_batching_only": round(walls["geo_single"] / max(walls["geo_batched"], 1e-9), 3),
        "held_out": total,
        "per_family": {f: summary([r for r in rows if r["family"] == f]) for f in ladder.MIX_FAMILIES
                       if any(r["family"] == f for r in rows)},
        "total_s": total_s,
        "rows": rows,
    }


def _tune_singles(models, seeds, targets, kw) -> None:
    """Fill the autotune cache for every shape class among models (runs that stop at a target),
    untimed: a two-round call of each instance tunes its class on a miss and hits the cache
    afterwards."""
    from drift_amd import gpu

    warm = dict(kw, n_rounds=2, record_history=False)
    for m, s, t in zip(models, seeds, targets):
        gpu.parallel_tempering_gpu(m, seed=s, target_energy=t, autotune=True, **warm)


def bench_ttt_sample(per_family: int = 8, cap: int = 100_000) -> dict:
    """Time to target on a sample of the held-out set, as a sub-result of the default line:
    per_family held-out instances of each family (N ~ 65536) spread evenly over its indices, every
    run stopping at the instance's committed best-known energy or after cap sweeps; (a) one call per instance on the
    geometric ladder on the autotuned engine (tuned untimed first), (b) one batched call on the
    committed network's learned ladders.  Same seeds and targets as the full evaluation
    (--workload ladder-c5 --stage eval), so each run equals that evaluation's run of the
    instance."""
    from drift_amd import gpu, ladder

    with open(os.path.join(ROOT, "tests", "golden", "c5_heldout_bestknown.json")) as f:
        bk = json.load(f)
    pf, n_target = bk["per_family"], bk["test_n"]
    rows = {r["index"]: r for r in bk["rows"]}
    net = ladder.LadderNet(seed=11)
    net.load(os.path.join(ROOT, "drift_amd", "ladder_net_c5.npz"))
    models, idx, targets, learned, ranges = [], [], [], [], []
    ks = [(j * pf) // per_family for j in range(per_family)]
    for fi, (fam, (T_min, T_max)) in enumerate(ladder.MIX_FAMILIES.items()):
        fam_models = [ladder.mix_instance(fam, 900_000 + 1000 * fi + k, n_target) for k in ks]
        xs = net.predict_ladder(ladder.instance_features(fam_models, T_min=T_min, T_max=T_max, R=32))
        for k, m, x in zip(ks, fam_models, xs):
            i = fi * pf + k
            if rows[i]["n"] != m.n:
                raise RuntimeError(f"held-out instance {i} does not match its best-known row")
            models.append(m)
            idx.append(i)
            targets.append(rows[i]["best_known_energy"])
            learned.append(ladder.betas_from_positions(x, T_min, T_max))
            ranges.append((fam, T_min, T_max))
    log(f"time to target: {len(models)} held-out instances, cap {cap} sweeps; tuning the engines")
    for m, i, t, (fam, T_min, T_max) in zip(models, idx, targets, ranges):
        _tune_singles([m], [4242 + i], [t], dict(n_replicas=32, T_min=T_min, T_max=T_max,
                                                 sweeps_per_round=1))
    log("time to target: one call each")
    t_single, single, engines = 0.0, [], {}
    for m, i, t, (fam, T_min, T_max) in zip(models, idx, targets, ranges):
        t0 = time.perf_counter()
        single.append(gpu.parallel_tempering_gpu(m, n_replicas=32, T_min=T_min, T_max=T_max,
                                                 n_rounds=cap, sweeps_per_round=1, seed=4242 + i,
                                                 target_energy=t, record_history=False,
                                                 autotune=True))
        r = single[-1]
        engines.setdefault(fam, set()).add(
            f"{r.engine}" + (f" {r.slots_per_group}x{r.parts_per_group}" if r.engine == "split" else ""))
        t_single += time.perf_counter() - t0
        log(f"time to target: instance {i} ({fam}) {'reached' if single[-1].best_E <= t else 'not reached'} "
            f"after {single[-1].rounds_run} sweeps")
    log("time to target: one batched call on the learned ladders")
    t0 = time.perf_counter()
    batched = gpu.parallel_tempering_gpu_batch(models, n_replicas=32, n_rounds=cap, sweeps_per_round=1,
                                               seeds=[4242 + i for i in idx], betas=learned,
                                               targets=targets)
    t_batched = time.perf_counter() - t0
    reached = lambda res: int(sum(r.best_E <= t + 1e-9 * abs(t) for r, t in zip(res, targets)))
    return {"metric": "time-to-target energy, learned vs geometric ladder (sample of the held-out set)",
            "value": round(t_single / max(t_batched, 1e-9), 3),
            "unit": "x less wall time (batched + learned vs one call per instance + geometric)",
            "higher_is_better": True, "steps": cap,
            "config": {"workload": f"{per_family} held-out instances of each family of the "
                                   f"learned-ladder evaluation (indices {ks} of {pf}; N ~ "
                                   f"{n_target}), R=32, best-known targets from "
                                   f"tests/golden/c5_heldout_bestknown.json",
                       "cap_sweeps": cap, "instances": len(models),
                       "batched_engine": batched[0].engine,
                       "single_engines": {f: sorted(v) for f, v in engines.items()}},
            "reached": {"geometric_one_call_each": reached(single),
                        "learned_batched": reached(batched), "of": len(models)},
            "wall_s": {"geometric_one_call_each": round(t_single, 3),
                       "learned_batched": round(t_batched, 3)},
            "full_evaluation": "profiles/bench_ladder_c5_eval.json (201 instances)"}


def bench_c3_ttt(args) -> dict:
    """C3's time-to-target report: the best energy a plain run of 10 x --steps sweeps finds (seed 1,
    geometric ladder) is the target; runs of at most --steps sweeps from another seed, on the
    geometric ladder and on the committed network's ladder, stop when they reach it.  Wall time is
    the library call's host time; device time is the sweep loop's."""
    from drift_amd import gpu, ladder

    model, R, beta, S, desc = workload("c3")
    T_min, T_max = 1.0 / beta[0], 1.0 / beta[-1]
    cap = args.steps
    log(f"c3-ttt: plain run of {10 * cap} sweeps for the target")
    t0 = time.perf_counter()
    long = gpu.parallel_tempering_gpu(model, n_replicas=R, betas=beta, n_rounds=10 * cap,
                                      sweeps_per_round=1, seed=1, record_history=False)
    long_wall = time.perf_counter() - t0
    target = long.best_E
    net = ladder.LadderNet(seed=0)
    net.load(os.path.join(ROOT, "drift_amd", "ladder_net_c5.npz"))
    x = net.predict_ladder(ladder.instance_features([model], T_min=T_min, T_max=T_max, R=R))[0]
    runs = {}
    for name, b in (("geometric", beta), ("learned", ladder.betas_from_positions(x, T_min, T_max))):
        log(f"c3-ttt: {name} ladder, at most {cap} sweeps")
        t0 = time.perf_counter()
        r = gpu.parallel_tempering_gpu(model, n_replicas=R, betas=b, n_rounds=cap, sweeps_per_round=1,
                                       seed=2, target_energy=target, record_history=True)
        wall = time.perf_counter() - t0
        runs[name] = {"reached": bool(r.best_E <= target), "rounds_run": r.rounds_run,
                      "wall_s": round(wall, 3), "device_s": round(r.seconds, 3),
                      "best_E": r.best_E, "gap_to_target": (r.best_E - target) / abs(target),
                      "within_1e-4_after": _reach(r.best_history[:max(r.rounds_run, 1)], target, 1e-4),
                      "within_1e-3_after": _reach(r.best_history[:max(r.rounds_run, 1)], target, 1e-3),
                      "engine": r.engine}
    return {"metric": "time to the best energy of a 10x longer plain run (C3)",
            "value": runs["geometric"]["wall_s"], "unit": "s (geometric ladder; the cap's time if "
            "not reached)", "higher_is_better": False, "n_gpus": 1, "steps": cap,
            "config": {"workload": desc, "target_run_sweeps": 10 * cap, "target_run_seed": 1,
                       "timed_seed": 2, "learned_power": float(np.log(x[8]) / np.log(8 / 31))},
            "target": target, "target_found_after_round": long.best_round,
            "target_run_wall_s": round(long_wall, 2), "runs": runs}


def bench_single_heldout(args) -> dict:
    """One instance at a time: the first held-out instance of each family of the learned-ladder
    mix (N ~ 65536, R = 32, geometric ladder, one sweep per round, --steps rounds) on the default
    engine, the per-launch engine, one-slot split groups in 4 and 8 parts and (unless
    --no-autotune) the autotuner's choice; --engine E keeps only the configurations named E
    (e.g. "split 1x8") or on engine E (a profiler command's way to leave out the graph-replayed
    engines).  Each value is the median of --repeats timed runs after
    one untimed run, in spin-flip attempts per ns; every configuration must end at the same best
    energy (one chain)."""
    from drift_amd import gpu, ladder

    fams = args.families.split(",") if args.families else list(ladder.MIX_FAMILIES)
    steps = args.steps
    configs = [("default", dict(engine="auto")), ("global", dict(engine="global")),
               ("split 1x4", dict(engine="split", slots_per_group=1, parts_per_group=4,
                                  block_threads=1024)),
               ("split 1x8", dict(engine="split", slots_per_group=1, parts_per_group=8,
                                  block_threads=1024))]
    if args.autotune:
        configs.append(("autotuned", dict(engine="auto", autotune=True)))
    if args.engine != "auto":   # --engine E: only the configurations named E or on engine E
        configs = [c for c in configs if args.engine in (c[0], c[1]["engine"])]
    out = {}
    for fi, (fam, (T_min, T_max)) in enumerate(ladder.MIX_FAMILIES.items()):
        if fam not in fams:
            continue
        model = ladder.mix_instance(fam, 900_000 + 1000 * fi, args.test_n)
        kw = dict(n_replicas=32, T_min=T_min, T_max=T_max, n_rounds=steps, sweeps_per_round=1,
                  seed=1, record_history=False, plain_launch=args.plain_launch)
        rows, best = {}, set()
        for name, extra in configs:
            log(f"single held-out: {fam} (N = {model.n}) on {name}")
            gpu.parallel_tempering_gpu(model, **dict(kw, **extra))
            runs = [gpu.parallel_tempering_gpu(model, **dict(kw, **extra))
                    for _ in range(max(args.repeats, 1))]
            secs = float(np.median([r.seconds for r in runs]))
            r = runs[0]
            best.add(r.best_E_units)
            rows[name] = {"attempts_per_ns": round(32 * steps * model.n / secs / 1e9, 2),
                          "engine": r.engine, "block_threads": r.block_threads,
                          "slots_per_group": r.slots_per_group,
                          "parts_per_group": r.parts_per_group}
        if len(best) != 1:
            raise RuntimeError(f"single held-out: engines disagree on {fam}")
        out[fam] = {"n": model.n, "engines": rows}
    return {"metric": "single-instance throughput by engine (held-out instances, N ~ 65536)",
            "unit": "attempts/ns", "higher_is_better": True, "n_gpus": 1, "steps": steps,
            "repeats": args.repeats, "data": "synthetic",
            "config": {"workload": "first held-out instance of each family of the learned-ladder "
                                   "mix, R=32, geometric ladder, one sweep per round"},
            "families": out}


def bench_parts_sweep(args) -> dict:
    """The single-instance engine choice across shapes: 3-regular N(0,1) graphs (N 8192 ..
    131072, four colour classes), 3-D +/-J glasses (L 20, 24, 40, 48; two classes) and sparse
    Erdos-Renyi graphs (five to seven classes), at R = 32 and 16, geometric ladder 0.2 .. 5.0,
    one sweep per round, --steps rounds: the default engine, the per-launch engine and one-slot
    split groups in 8 and 16 parts (1024 threads), each the median of --repeats timed runs after an
    untimed one, in attempts/ns.  Every configuration of a shape must end at the same best
    energy."""
    from drift_amd import gpu
    from drift_amd.builders.lattice import cubic_pmj_glass, er_sparse_weighted, random_regular_gaussian
    from drift_amd.solvers.parallel_tempering import geometric_ladder, ladder_betas

    shapes = [("3-regular", n, lambda n=n: random_regular_gaussian(n, seed=3))
              for n in (8192, 16384, 32768, 65536, 131072)]
    shapes += [("3-D glass", L ** 3, lambda L=L: cubic_pmj_glass(L, 1)) for L in (20, 24, 40, 48)]
    shapes += [("ER degree 3", 65536, lambda: er_sparse_weighted(65536, 3.0, 2)),
               ("ER degree 6", 65536, lambda: er_sparse_weighted(65536, 6.0, 2)),
               ("ER degree 3", 16384, lambda: er_sparse_weighted(16384, 3.0, 2))]
    configs = [("default", dict(engine="auto")), ("global", dict(engine="global")),
               ("split 1x8", dict(engine="split", slots_per_group=1, parts_per_group=8,
                                  block_threads=1024)),
               ("split 1x16", dict(engine="split", slots_per_group=1, parts_per_group=16,
                                   block_threads=1024))]
    rows = []
    for R in (32, 16):
        beta = ladder_betas(geometric_ladder(R, 0.2, 5.0))
        for fam, n, make in shapes:
            model = make()
            row = {"family": fam, "n": n, "R": R}
            best = set()
            for name, kw in configs:
                if "parts_per_group" in kw and R * kw["parts_per_group"] > 256:
                    continue   # more workgroups than compute units
                run = dict(n_replicas=R, n_rounds=args.steps, sweeps_per_round=1, seed=1,
                           betas=beta, record_history=False, **kw)
                gpu.parallel_tempering_gpu(model, **dict(run, n_rounds=2))
                res = [gpu.parallel_tempering_gpu(model, **run) for _ in range(max(args.repeats, 1))]
                secs = float(np.median([r.seconds for r in res]))
                best.add(res[0].best_E_units)
                row[name] = round(R * args.steps * n / secs / 1e9, 2)
                if name == "default":
                    row["default_engine"] = res[0].engine + (
                        f" 1x{res[0].parts_per_group}" if res[0].engine == "split" else "")
            if len(best) != 1:
                raise RuntimeError(f"parts sweep: engines disagree on {fam} n={n} R={R}")
            log(f"parts sweep: {row}")
            rows.append(row)
    return {"metric": "single-instance throughput by engine across shapes", "unit": "attempts/ns",
            "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "repeats": args.repeats,
            "data": "synthetic", "rows": rows}


def bench_ladder_c5(args) -> int:
    stage = {"train": c5_train, "bestknown": c5_bestknown, "eval": c5_eval,
             "eval-merge": c5_eval_merge, "eval-batched": c5_eval_batched}[args.stage]
    line = stage(args)
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=5, help="timed runs; the value is their median")
    ap.add_argument("--workload", default="all",
                    choices=("all", "c4", "c2", "c3", "c1", "published", "c5-batch", "c3-ttt",
                             "ladder-c5", "ttt-sample", "single-heldout", "parts-sweep",
                             "single3"))
    ap.add_argument("--sub-steps", type=int, default=0,
                    help="steps of the sub-result workloads (0: their defaults)")
    ap.add_argument("--engine", default="auto")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="take the library's default engine and launch shape instead of timing "
                         "every candidate first (the tuning runs before the timed runs, and the "
                         "default stays unless a candidate is more than 3%% faster)")
    ap.add_argument("--plain-launch", action="store_true",
                    help="ordinary dispatch of the persistent engines (for profiler commands)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-target", action="store_true",
                    help="leave the time-to-target sample out of the default line")
    ap.add_argument("--instances", type=int, default=4096, help="batch size of the c4 workload")
    ap.add_argument("--batch-family", default="rrg", choices=("rrg", "er"),
                    help="c5-batch workload: 3-regular float graphs or the held-out sparse ER graphs")
    ap.add_argument("--batch-count", type=int, default=64, help="c5-batch workload: instances")
    ap.add_argument("--train-per-family", type=int, default=667)
    ap.add_argument("--stage", default="train",
                    choices=("train", "bestknown", "eval", "eval-merge", "eval-batched"),
                    help="ladder-c5 workload: train the network, record the held-out instances' "
                         "best-known energies, or measure time to them")
    ap.add_argument("--net", default=os.path.join(ROOT, "drift_amd", "ladder_net_c5.npz"),
                    help="ladder-c5 network file")
    ap.add_argument("--label-budget", type=int, default=5000,
                    help="ladder-c5: sweeps of each labelling run")
    ap.add_argument("--cap", type=int, default=5000,
                    help="ladder-c5: sweeps a held-out run may take to reach its target (the "
                         "best-known runs take ten times as many)")
    ap.add_argument("--bestknown", default=os.path.join(ROOT, "tests", "golden",
                                                        "c5_heldout_bestknown.json"))
    ap.add_argument("--out", default=os.path.join(ROOT, "c5_out"),
                    help="ladder-c5: directory of the stage files (labels, best-known rows, "
                         "evaluation parts)")
    ap.add_argument("--test-n", type=int, default=65536, help="ladder-c5 held-out size")
    ap.add_argument("--test-per-family", type=int, default=67)
    ap.add_argument("--families", default="",
                    help="ladder-c5 bestknown / eval: comma-separated families to run (default "
                         "all); bestknown keeps the other families' rows already in --out, eval "
                         "writes the families' rows to --out for --stage eval-merge")
    ap.add_argument("--parts", default="",
                    help="ladder-c5 eval-merge / eval-batched: comma-separated part files of "
                         "--stage eval")
    ap.add_argument("--learned-only", action="store_true",
                    help="ladder-c5 eval-batched: rerun the learned batched leg only, taking the "
                         "geometric batched leg from the part files")
    ap.add_argument("--slots-per-group", type=int, default=0,
                    help="ladder-c5 eval-batched: the split engine's slots per group in the "
                         "batched calls (0: the library's choice)")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many bytes (rocprof counter scaling)")
    args = ap.parse_args()
    args.autotune = not args.no_autotune
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this engine has no "
              "multi-GPU path (batched instances would shard across GPUs with no communication, "
              "one instance's ladder by contiguous rung blocks); one GPU is measured",
              file=sys.stderr)
        return 2

    if args.workload == "ladder-c5" and args.stage == "eval-merge":   # host-only
        print(json.dumps(c5_eval_merge(args)), flush=True)
        return 0

    from drift_amd import gpu

    if not gpu.gpu_available():
        print("bench.py: GPU engine not available (no gfx950 device or library not built)",
              file=sys.stderr)
        return 1
    if args.workload == "ladder-c5":
        return bench_ladder_c5(args)
    if args.workload == "c3-ttt":
        print(json.dumps(bench_c3_ttt(args)), flush=True)
        return 0
    if args.workload == "ttt-sample":
        print(json.dumps(bench_ttt_sample()), flush=True)
        return 0
    if args.workload == "single-heldout":
        print(json.dumps(bench_single_heldout(args)), flush=True)
        return 0
    if args.workload == "parts-sweep":
        print(json.dumps(bench_parts_sweep(args)), flush=True)
        return 0
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
