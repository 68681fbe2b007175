/*
 * drift_pt.h -- C ABI of libdriftpt.so, the gfx950 parallel-tempering engine.
 *
 * It replaces the reference's process boundary: drift/gpu.py:83-106 wrote problem.bin, spawned
 * cuda/ising_pt and parsed result.bin plus a stderr line; cuda/ising_pt.cu:159-243 was that
 * executable.  Here the caller hands plain host arrays to one synchronous call and receives the
 * results in caller-allocated host arrays.  The library copies inputs to device memory, never keeps
 * a caller pointer after returning, and never calls exit().
 *
 * Conventions (drift/ising.py:25-68): spins are +1/-1, E(s) = -1/2 s^T J s - h^T s, the coupling
 * matrix is symmetric with a zero diagonal and is passed as CSR holding both directions of every
 * bond.  Rung 0 is the coldest rung (largest beta), as in the reference's ascending T ladder
 * (drift/solvers/parallel_tempering.py:29-37).
 *
 * All entry points return 0 on success and a negative drift_pt_status otherwise; the message is
 * then available from drift_pt_last_error() (thread-local).
 */
#ifndef DRIFT_PT_H
#define DRIFT_PT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRIFT_PT_ABI_VERSION 6

enum drift_pt_status {
    DRIFT_PT_OK = 0,
    DRIFT_PT_EINVAL = -1,   /* malformed argument or CSR */
    DRIFT_PT_ECOLOR = -2,   /* colouring is not a partition or two same-colour spins are adjacent */
    DRIFT_PT_ENOMEM = -3,   /* device or host allocation failed */
    DRIFT_PT_EHIP = -4,     /* a HIP runtime call failed */
    DRIFT_PT_ENODEV = -5,   /* no gfx950 device visible */
    DRIFT_PT_ERANGE = -6    /* values outside what the exact integer path can represent */
};

/* coupling storage (cuda/ising_pt.cu:44 used float weights only) */
enum drift_pt_weight_type {
    DRIFT_PT_W_I8 = 1,      /* int8 weights, int32 field: exact integer path */
    DRIFT_PT_W_I32 = 2,     /* int32 weights, int32 field: exact integer path */
    DRIFT_PT_W_F32 = 3      /* float weights, float field */
};

enum drift_pt_engine {
    DRIFT_PT_ENGINE_AUTO = 0,   /* the library's choice: LDS for small instances; lattices on the
                                   slab engines where they fit; one instance of n >= 8000 with at
                                   most four colour classes and fewer replicas than compute units
                                   on SPLIT with one-slot groups in parts; otherwise COOP for
                                   R n <= 2^17, else LATTICE or GLOBAL (DRIFT_PT_AUTOTUNE times
                                   the candidates instead) */
    DRIFT_PT_ENGINE_LDS = 1,    /* whole instance (all replicas) in one workgroup's LDS, persistent */
    DRIFT_PT_ENGINE_GLOBAL = 2, /* replica-interleaved spins in HBM, one launch per colour step */
    DRIFT_PT_ENGINE_LATTICE = 3,/* the global engine's periodic-lattice kernel: implicit neighbours,
                                   one byte of bond signs per site (+/-1 couplings, no field) */
    DRIFT_PT_ENGINE_SLAB = 4,   /* periodic lattices: one launch per sweep, a workgroup per slab of
                                   planes of one replica, halo planes recomputed in LDS */
    DRIFT_PT_ENGINE_COOP = 5,   /* CSR instances too large for one workgroup: one cooperative launch
                                   runs every round; each workgroup keeps its share of the rows in
                                   LDS, colour steps are separated by grid barriers */
    DRIFT_PT_ENGINE_PSLAB = 6,  /* the slab engine's lattices in one cooperative launch: slabs stay
                                   in LDS for the whole run, halo planes move through HBM between
                                   sweeps, one grid barrier per sweep.  The first run of a lattice
                                   shape in a process times the barrier's two energy paths and
                                   keeps the faster (environment DRIFT_PT_PSLAB_ENERGY_PATH=atomic
                                   or =words fixes it); results do not depend on the path */
    DRIFT_PT_ENGINE_SPLIT = 7   /* instances whose replicas do not fit one workgroup's LDS (and
                                   batches of them): the replica slots are split into groups of
                                   params.slots_per_group, each group's spins in one workgroup's
                                   LDS for the whole run; after each round's sweeps the groups
                                   trade slot energies through round-tagged mailboxes in HBM and
                                   all take the same exchange decisions.  One persistent launch;
                                   every group of an instance is resident at once.  With
                                   params.parts_per_group = P > 1 each group's positions are
                                   split into P parts, one workgroup each (every one holding the
                                   group's whole state): after each colour step the parts trade
                                   that class's new spins as step-tagged bit words, so one large
                                   instance with few replicas fills the chip */
};

enum drift_pt_flags {
    DRIFT_PT_RECORD_HISTORY = 1, /* fill result.best_history */
    DRIFT_PT_RECORD_FINAL = 2,   /* fill result.final_spins */
    DRIFT_PT_SKIP_VALIDATION = 4, /* trust CSR symmetry/colouring (the colouring is still used) */
    DRIFT_PT_RECORD_TRACE = 8,    /* fill result.rung_energy_trace */
    DRIFT_PT_TIME_KERNELS = 16,   /* time each sweep-kernel launch (up to 4096) with device events */
    DRIFT_PT_NO_GRAPH = 32,       /* launch every round directly instead of replaying a hipGraph */
    DRIFT_PT_AUTOTUNE = 64,       /* with DRIFT_PT_ENGINE_AUTO: take the engine and launch shape
                                     from the autotune cache, timing the candidates on a miss
                                     (with DRIFT_PT_STOP_AT_TARGET: among the engines that can
                                     stop at a target) */
    DRIFT_PT_PLAIN_LAUNCH = 128,  /* persistent engines (coop, pslab): an ordinary dispatch of the
                                     occupancy-sized grid instead of a cooperative launch.  All
                                     workgroups are resident when nothing else runs on the device;
                                     otherwise the bounded grid barrier fails the call after its
                                     time bound.  For profilers that cannot follow cooperative
                                     launches.  (Environment DRIFT_PT_BARRIER_TIMEOUT_MS sets the
                                     bound of one barrier wait; default 20000.) */
    DRIFT_PT_STOP_AT_TARGET = 256,/* stop after the first round end whose best energy is at or
                                     below params.target_energy (within 1e-9 relative); the result
                                     is then that of a run of result.rounds_run rounds.  Engines
                                     LDS, SPLIT, GLOBAL and LATTICE (AUTO picks among them); the
                                     others refuse it. */
    DRIFT_PT_COUPLINGS_LDS = 512  /* SPLIT engine: each workgroup also stages its instance's
                                     couplings in LDS (row pointers, 16-bit records, field), so
                                     every neighbour gather is an LDS read.  Needs int8 couplings
                                     with |w| <= 7 and n <= 4096; in a batch, every instance of
                                     the split launch must ask for it */
};

/* One Ising instance.  Replaces the problem.bin layout of cuda/ising_pt.cu:40-45 and
 * drift/gpu.py:83-95. */
typedef struct drift_pt_problem {
    int32_t n;                  /* spins */
    int32_t nnz;                /* CSR entries, both directions of each bond */
    const int32_t* row_ptr;     /* n + 1 */
    const int32_t* col_idx;     /* nnz, no self loops */
    int32_t weight_type;        /* enum drift_pt_weight_type */
    int32_t n_colors;           /* k; 0 = colour greedily inside the library */
    const void* weight;         /* nnz entries of weight_type */
    const void* h;              /* n entries: int32 for integer types, float for F32; NULL = 0 */
    double energy_scale;        /* energy = scale * integer energy (integer types; 1.0 if unsure) */
    const int32_t* color_ptr;   /* k + 1 (ignored when n_colors == 0) */
    const int32_t* color_spins; /* n, spins grouped by colour */
} drift_pt_problem;

/* Run parameters (drift/gpu.py:43-52 keyword arguments). */
typedef struct drift_pt_params {
    uint64_t seed;
    int32_t n_replicas;         /* R */
    int32_t n_rounds;
    int32_t sweeps_per_round;
    int32_t engine;             /* enum drift_pt_engine */
    double T_min, T_max;        /* geometric ladder when beta == NULL */
    const double* beta;         /* optional R inverse temperatures, rung 0 coldest */
    int32_t flags;              /* enum drift_pt_flags */
    int32_t block_threads;      /* 0 = chosen by the tuner */
    double target_energy;       /* with DRIFT_PT_STOP_AT_TARGET: the energy to stop at */
    int32_t slots_per_group;    /* SPLIT engine: replica slots per workgroup, a power of two up
                                   to 64; 0 = chosen by the library (or the tuner) */
    int32_t parts_per_group;    /* SPLIT engine: workgroups sharing each group's positions, 1 to
                                   64 (groups x parts <= 1024 per instance); 0 = chosen by the
                                   library (or the tuner) */
} drift_pt_params;

/* Results, all arrays caller-allocated (NULL = not wanted, except best_s). */
typedef struct drift_pt_result {
    int8_t* best_s;             /* n: lowest-energy configuration seen (original spin order) */
    double best_E;              /* its energy */
    int64_t best_E_units;       /* exact integer units (integer types) or 2^-36 units (F32) */
    int32_t best_round;         /* 0 = initial states, r + 1 = after round r */
    int32_t best_rung;
    int32_t best_slot;
    int32_t engine_used;
    int64_t* swap_tried;        /* R - 1, per lower rung */
    int64_t* swap_accepted;     /* R - 1 */
    int32_t* round_trips;       /* R, per replica slot: coldest -> hottest -> coldest trips */
    int64_t* rung_up;           /* R, per rung: rounds spent there by replicas last at the cold end */
    int64_t* rung_down;         /* R, per rung: rounds spent there by replicas last at the hot end */
    int32_t* rung_of;           /* R, final rung of each slot */
    int64_t* final_E_units;     /* R, final energy of each slot in units */
    int8_t* final_spins;        /* R * n slot-major (needs DRIFT_PT_RECORD_FINAL) */
    double* best_history;       /* n_rounds (needs DRIFT_PT_RECORD_HISTORY) */
    double* rung_energy_trace;  /* n_rounds * R: energy at each rung after each round's sweeps
                                   (needs DRIFT_PT_RECORD_TRACE) */
    double seconds;             /* device time of the sweep + exchange loop (cuda/ising_pt.cu:213-222) */
    double attempts;            /* R * n_rounds * sweeps_per_round * n (cuda/ising_pt.cu:222) */
    double kernel_seconds;      /* summed device time of the timed sweep-kernel launches */
    double kernel_attempts;     /* spin-flip attempts those launches performed */
    int64_t kernel_launches;    /* launches timed (DRIFT_PT_TIME_KERNELS), else 1 for the whole loop */
    double other_kernel_seconds;    /* DRIFT_PT_TIME_KERNELS: summed device time of the timed
                                       non-sweep launches (round ends of the per-launch engines) */
    int64_t other_kernel_launches;  /* and their number (0 when the sweep kernel does everything) */
    int32_t rounds_run;         /* rounds run: n_rounds, or fewer when the target was reached */
    int32_t block_threads_used; /* workgroup size of the sweep kernel */
    int32_t slots_per_group_used; /* replica slots per workgroup (the SPLIT engine's groups; the
                                     padded slot count on the LDS engine; 0 on the others) */
    int32_t couplings_in_lds_used; /* 1 when the couplings were staged in LDS */
    int32_t parts_per_group_used;  /* SPLIT: workgroups per replica group (1 = one each); 0 on
                                      the other engines */
} drift_pt_result;

/* Timed candidates of one autotune call (drift_pt_autotune). */
#define DRIFT_PT_TUNE_MAX 64
typedef struct drift_pt_tuning {
    int32_t engine;                          /* the chosen candidate */
    int32_t block_threads;
    int32_t slots_per_group;                 /* SPLIT: replica slots per workgroup, else 0 */
    int32_t couplings_in_lds;                /* SPLIT: 1 with the couplings staged in LDS */
    int32_t parts_per_group;                 /* SPLIT: workgroups per replica group, else 0 */
    int32_t trial_rounds;                    /* rounds each timed trial ran */
    int32_t n_candidates;
    int32_t cand_engine[DRIFT_PT_TUNE_MAX];
    int32_t cand_block[DRIFT_PT_TUNE_MAX];
    int32_t cand_group[DRIFT_PT_TUNE_MAX];
    int32_t cand_lds_couplings[DRIFT_PT_TUNE_MAX];
    int32_t cand_parts[DRIFT_PT_TUNE_MAX];
    double cand_seconds[DRIFT_PT_TUNE_MAX];  /* device time of trial_rounds rounds (a run of
                                                2 * trial_rounds minus a run of trial_rounds, so
                                                fixed costs cancel); < 0 if it could not run */
} drift_pt_tuning;

/* ABI version of the loaded library (DRIFT_PT_ABI_VERSION). */
int drift_pt_abi_version(void);

/* 1 if a gfx950 device is visible, else 0.  Never fails.  Replaces the binary-exists probe of
 * drift/gpu.py:38-40. */
int drift_pt_device_available(void);

/* Message of the last failing call on this thread ("" if none). */
const char* drift_pt_last_error(void);

/* Greedy colouring of a CSR graph (drift/gpu.py:122-145): vertices in index order take the smallest
 * colour unused by earlier neighbours; color_spins is the stable sort by colour.  color_ptr must hold
 * n + 1 entries.  Returns k (>= 1) or a negative status. */
int drift_pt_greedy_coloring(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                             int32_t* color_ptr, int32_t* color_spins);

/* One parallel-tempering run on the GPU.  Replaces cuda/ising_pt.cu:159-243 (main) and the
 * subprocess round trip of drift/gpu.py:83-106. */
int drift_pt_run(const drift_pt_problem* problem, const drift_pt_params* params,
                 drift_pt_result* result);

/* Many independent instances; each result equals what drift_pt_run gives for that instance and
 * its params alone.  Instances whose replicas fit one workgroup's LDS run in one persistent launch
 * (a work queue, largest first); larger ones whose single replica fits run in a second persistent
 * launch on the SPLIT engine when their workgroups (replica groups, times params.parts_per_group
 * where given) number at least one per compute unit; the rest run on the per-launch engine,
 * several at a time on streams of their own so their launches overlap.  With DRIFT_PT_AUTOTUNE on
 * any instance, each persistent launch's shape (block size; for the split launch also slots per
 * group and couplings in LDS) is timed on the batch first and cached per batch shape class;
 * params.block_threads / slots_per_group / parts_per_group fix it instead.  No reference
 * counterpart: the reference solves one instance per process (drift/gpu.py:97). */
int drift_pt_run_batch(const drift_pt_problem* problems, const drift_pt_params* params,
                       int32_t n_instances, drift_pt_result* results);

/* Size of the device region drift_pt_run_batch_noalloc works in: 64 MiB per device, about 900
 * instances of 1000 spins with 16 replicas (each needs about 70 KB: its packed CSR, padded
 * first-chunk records, field, ladder, statistics block and best state). */
#define DRIFT_PT_NOALLOC_BYTES (64u << 20)

/* drift_pt_run_batch without any device allocation, for callers that cannot allocate (code
 * between graph captures, tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * same results bit for bit.  Every device buffer of the call -- the packed instances, their
 * descriptors, the work queue and the outputs -- is carved from a region of
 * DRIFT_PT_NOALLOC_BYTES that the library holds on each device for the life of the process; the
 * call creates no stream or event (it runs on the null stream, and result.seconds is the host's
 * clock around the launch).  A batch needing more than the region fails with DRIFT_PT_ERANGE.
 * Calls from several threads on one device are serialised, because they share the region; the
 * region keeps the last call's contents.  No reference counterpart. */
int drift_pt_run_batch_noalloc(const drift_pt_problem* problems, const drift_pt_params* params,
                               int32_t n_instances, drift_pt_result* results);

/* Autotuner.  Times every engine and launch shape that can run this instance (LDS, split, slab,
 * lattice, global, coop, pslab; 256 / 512 / 1024 threads; for the split engine every power-of-two
 * number of replica slots per workgroup whose spins fit LDS, with the couplings in HBM and, where
 * they qualify, in LDS) for trial_rounds rounds (0: about
 * 2.5e8 attempts), in two passes of opposite order with the faster of each candidate's two trials
 * kept, stores the choice in a process-wide cache keyed by the instance's shape class (weight
 * type, replica slots, lattice shape, size and degree buckets, sweeps per round), and reports
 * every candidate's time.  The engine the library takes without tuning stays the choice unless
 * another candidate is more than 3% faster, so trial noise does not move it.  All candidates run
 * the same Markov chain, so the choice changes speed only.
 * No reference counterpart: the reference has one kernel with a fixed launch shape
 * (cuda/ising_pt.cu:196-211). */
int drift_pt_autotune(const drift_pt_problem* problem, const drift_pt_params* params,
                      int32_t trial_rounds, drift_pt_tuning* out);

/* Empties the autotune cache. */
void drift_pt_autotune_clear(void);

#ifdef __cplusplus
}
#endif
#endif
