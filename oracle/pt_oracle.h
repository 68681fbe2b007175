/*
 * pt_oracle.h -- CPU restatement of the parallel-tempering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in drift_amd/ links, loads or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
 *
 * What it restates (reference files, cited relative to the reference root):
 *   - IsingModel energy / flip convention        drift/ising.py:50-68
 *   - geometric ladder, beta = 1/max(T,1e-12)    drift/solvers/parallel_tempering.py:29-37,72-73
 *   - colour-ordered Metropolis sweeps           cuda/ising_pt.cu:109-121
 *   - accept iff dE <= 0 or u < exp(-beta dE)   cuda/ising_pt.cu:118, parallel_tempering.py:96
 *   - alternating-parity adjacent-rung exchange  drift/solvers/parallel_tempering.py:104-113
 *   - best tracking at round granularity         parallel_tempering.py:79-82,99-102
 *   - greedy colouring                           drift/gpu.py:122-145
 *   - exact enumeration, bit b -> spin 1-2b      drift/solvers/exact.py:20-49
 * and the engine's own counter-based stream (Philox4x32-10, Salmon et al., SC'11), which the
 * reference does not have: the reference draws from cuRAND / NumPy, so its streams are not
 * reproducible here; the stream contract below is this engine's, and the GPU engine must be
 * bit-identical to this file under it.
 *
 * Stream contract (key = seed: k0 = low 32 bits, k1 = high 32 bits):
 *   init spin i of slot a   : philox(ctr = (i>>2, a, 0, 1))[i&3] & 1 ? +1 : -1
 *   Metropolis, position p  : u = philox(ctr = (p>>2, a, t, 2))[p&3], t = round*S + sweep
 *   exchange, pair (r, r+1) : u = philox(ctr = (r>>2, 0, round, 3))[r&3]
 * Positions p index the concatenated colour lists (color_spins).
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_W_I8 = 1, OR_W_I32 = 2, OR_W_F32 = 3 };

typedef struct or_problem {
    int32_t n, nnz;
    const int32_t *row_ptr, *col_idx;
    int32_t weight_type;          /* OR_W_I8 / OR_W_I32 / OR_W_F32 */
    const void *weight;           /* nnz entries */
    const void *h;                /* n entries, int32 (integer types) or float; NULL = 0 */
    double energy_scale;          /* integer energy units -> energy */
    int32_t n_colors;
    const int32_t *color_ptr, *color_spins;
} or_problem;

typedef struct or_params {
    uint64_t seed;
    int32_t R, n_rounds, sweeps_per_round, n_threads;
    const double *beta;           /* R entries, rung 0 coldest */
} or_params;

typedef struct or_result {
    int8_t *best_s;               /* n */
    int64_t best_units;
    double best_E;
    int32_t best_round, best_rung, best_slot, pad;
    int64_t *swap_tried, *swap_acc;   /* R-1 */
    int32_t *round_trips;             /* R, per slot */
    int64_t *rung_up, *rung_down;     /* R, per rung */
    int32_t *rung_of;                 /* R, per slot */
    int64_t *final_units;             /* R, per slot */
    int8_t *final_spins;              /* R*n slot-major, original order */
    double *best_history;             /* n_rounds */
    double *rung_trace;               /* n_rounds * R: energy at each rung after the sweeps */
    double seconds;
} or_result;

void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_exp(double x);
float or_expf(float x);
uint32_t or_thr_int(double beta_scaled, int32_t q);
int32_t or_greedy_coloring(int32_t n, const int32_t *row_ptr, const int32_t *col_idx,
                           int32_t *color_ptr, int32_t *color_spins);
int32_t or_pt_run(const or_problem *p, const or_params *prm, or_result *res);
int32_t or_exact(int32_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *w,
                 const double *h, int32_t n_beta, const double *betas, double *e_min,
                 int64_t *idx_min, double *mean_E, double *mean_E2);

#ifdef __cplusplus
}
#endif
#endif
