/*
 * pt_oracle.c -- CPU restatement of the parallel-tempering path (see pt_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared against.  Built with
 * -ffp-contract=off so every float/double operation rounds exactly once, in program order.
 */
#include "pt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): 10 rounds of two 32x32->64 multiplies with multipliers 0xD2511F53 / 0xCD9E8D57 and a
 * Weyl key schedule 0x9E3779B9 / 0xBB67AE85.
 * ---------------------------------------------------------------------------------------- */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t m0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t m1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)m1;
        uint32_t n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)m0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint32_t philox_word(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                   uint64_t seed, int w) {
    uint32_t ctr[4] = {c0, c1, c2, c3}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    or_philox(ctr, key, o);
    return o[w & 3];
}

/* ------------------------------------------------------------------------------------------
 * Deterministic exponentials.  Cody-Waite reduction x = n ln2 + r with a split ln2 whose high
 * part has trailing zero bits (n*ln2_hi exact), then a Taylor polynomial in Horner form, then an
 * exact scaling by 2^n built from exponent bits.  The engine uses the same algorithm so that its
 * acceptance decisions are bit-identical; accuracy against libm is checked by the tests.
 * ---------------------------------------------------------------------------------------- */
double or_exp(double x) {
    if (!(x > -700.0)) return 0.0;            /* also NaN -> 0 */
    if (x > 700.0) x = 700.0;
    double nf = floor(x * 0x1.71547652b82fep+0 + 0.5);
    double r = (x - nf * 0x1.62e42fee00000p-1) - nf * 0x1.a39ef35793c76p-33;
    double p = 0x1.6124613a86d09p-33;         /* 1/13! */
    p = p * r + 0x1.1eed8eff8d898p-29;        /* 1/12! */
    p = p * r + 0x1.ae64567f544e4p-26;
    p = p * r + 0x1.27e4fb7789f5cp-22;
    p = p * r + 0x1.71de3a556c734p-19;
    p = p * r + 0x1.a01a01a01a01ap-16;
    p = p * r + 0x1.a01a01a01a01ap-13;
    p = p * r + 0x1.6c16c16c16c17p-10;
    p = p * r + 0x1.1111111111111p-7;
    p = p * r + 0x1.5555555555555p-5;
    p = p * r + 0x1.5555555555555p-3;
    p = p * r + 0x1.0000000000000p-1;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int n = (int)nf;
    uint64_t bits = (uint64_t)(int64_t)(n + 1023) << 52;
    double scale;
    memcpy(&scale, &bits, sizeof scale);
    return p * scale;
}

float or_expf(float x) {
    if (!(x > -87.0f)) return 0.0f;
    if (x > 87.0f) x = 87.0f;
    float nf = floorf(x * 0x1.715476p+0f + 0.5f);
    float r = (x - nf * 0x1.63p-1f) - nf * -0x1.bd0106p-13f;
    float p = 0x1.a01a02p-13f;                /* 1/7! */
    p = p * r + 0x1.6c16c2p-10f;
    p = p * r + 0x1.111112p-7f;
    p = p * r + 0x1.555556p-5f;
    p = p * r + 0x1.555556p-3f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int n = (int)nf;
    uint32_t bits = (uint32_t)(n + 127) << 23;
    float scale;
    memcpy(&scale, &bits, sizeof scale);
    return p * scale;
}

/* Integer-coupling Metropolis threshold: accept an uphill move of dE = 2 q (integer units)
 * iff u32 < floor(2^32 exp(-2 beta_scaled q)). */
uint32_t or_thr_int(double beta_scaled, int32_t q) {
    double x = -2.0 * beta_scaled * (double)q;
    double v = or_exp(x) * 4294967296.0;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

static int metro_float(float beta, float q, uint32_t u) {
    if (q <= 0.0f) return 1;
    float dE = 2.0f * q;
    float x = -(beta * dE);
    float p = or_expf(x);
    uint32_t thr = p >= 1.0f ? 0xFFFFFFFFu : (uint32_t)(p * 4294967296.0f);
    return u < thr;
}

static int exch_accept(double b_lo, double b_hi, double e_lo, double e_hi, uint32_t u) {
    double d = (b_lo - b_hi) * (e_lo - e_hi);
    if (d >= 0.0) return 1;
    return (double)u < or_exp(d) * 4294967296.0;
}

/* v in 2^-36 units, rounded to nearest with ties to even (exact for |v 2^36| < 2^51): the exact
 * product plus 1.5 * 2^52 keeps the integer in the low mantissa bits. */
static inline int64_t fix36d(double v) {
    union { double d; int64_t i; } c;
    c.d = v * 68719476736.0 + 6755399441055744.0;
    return c.i - (int64_t)0x4338000000000000ll;
}

/* ------------------------------------------------------------------------------------------
 * Greedy colouring, following drift/gpu.py:122-145: visit 0..n-1, give each vertex the smallest
 * colour no already-coloured neighbour holds; color_spins is the stable sort of vertices by
 * colour and color_ptr its offsets.  color_ptr must hold n+1 entries.  Returns k.
 * ---------------------------------------------------------------------------------------- */
int32_t or_greedy_coloring(int32_t n, const int32_t *row_ptr, const int32_t *col_idx,
                           int32_t *color_ptr, int32_t *color_spins) {
    if (n <= 0) { color_ptr[0] = 0; color_ptr[1] = 0; return 1; }
    int32_t *color = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t maxdeg = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t d = row_ptr[i + 1] - row_ptr[i];
        if (d > maxdeg) maxdeg = d;
        color[i] = -1;
    }
    char *used = (char *)calloc((size_t)maxdeg + 2, 1);
    int32_t k = 1;
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t) {
            int32_t c = color[col_idx[t]];
            if (c >= 0 && c <= maxdeg) used[c] = 1;
        }
        int32_t ci = 0;
        while (used[ci]) ++ci;
        color[i] = ci;
        if (ci + 1 > k) k = ci + 1;
        for (int32_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t) {
            int32_t c = color[col_idx[t]];
            if (c >= 0 && c <= maxdeg) used[c] = 0;
        }
    }
    for (int32_t c = 0; c <= k; ++c) color_ptr[c] = 0;
    for (int32_t i = 0; i < n; ++i) color_ptr[color[i] + 1]++;
    for (int32_t c = 0; c < k; ++c) color_ptr[c + 1] += color_ptr[c];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    for (int32_t c = 0; c < k; ++c) fill[c] = color_ptr[c];
    for (int32_t i = 0; i < n; ++i) color_spins[fill[color[i]]++] = i;
    free(fill);
    free(used);
    free(color);
    return k;
}

/* ------------------------------------------------------------------------------------------
 * Parallel tempering under the engine's stream contract.
 * ---------------------------------------------------------------------------------------- */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int32_t or_pt_run(const or_problem *P, const or_params *prm, or_result *res) {
    const int32_t n = P->n, R = prm->R, S = prm->sweeps_per_round;
    if (n <= 0 || R <= 0 || S < 0 || prm->n_rounds < 0 || !prm->beta) return -1;
    const int is_float = P->weight_type == OR_W_F32;
    const double unit = is_float ? 0x1p-36 : P->energy_scale;
    const int32_t *rp = P->row_ptr, *ci = P->col_idx;
    const int8_t *w8 = (const int8_t *)P->weight;
    const int32_t *w32 = (const int32_t *)P->weight;
    const float *wf = (const float *)P->weight;
    const int32_t *hi = (const int32_t *)P->h;
    const float *hf = (const float *)P->h;

    /* integer threshold table [rung][q-1], q up to the largest local field */
    int32_t qmax = 1;
    if (!is_float) {
        for (int32_t i = 0; i < n; ++i) {
            int64_t s = hi ? llabs((long long)hi[i]) : 0;
            for (int32_t t = rp[i]; t < rp[i + 1]; ++t)
                s += P->weight_type == OR_W_I8 ? abs(w8[t]) : llabs((long long)w32[t]);
            if (s > qmax) qmax = (int32_t)s;
        }
    }
    uint32_t *thr = NULL;
    if (!is_float) {
        thr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)R * (size_t)qmax);
        for (int32_t r = 0; r < R; ++r)
            for (int32_t q = 1; q <= qmax; ++q)
                thr[(size_t)r * qmax + (q - 1)] = or_thr_int(prm->beta[r] * P->energy_scale, q);
    }

    int8_t *s = (int8_t *)malloc((size_t)R * (size_t)n);
    int64_t *E = (int64_t *)malloc(sizeof(int64_t) * (size_t)R);
    int32_t *rung_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    int32_t *rep_at = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    int32_t *label = (int32_t *)calloc((size_t)R, sizeof(int32_t));

    /* init spins and energies (drift/ising.py:50-53 convention) */
    for (int32_t a = 0; a < R; ++a) {
        int8_t *sa = s + (size_t)a * n;
        for (int32_t i = 0; i < n; ++i)
            sa[i] = (philox_word((uint32_t)(i >> 2), (uint32_t)a, 0u, 1u, prm->seed, i) & 1u) ? 1 : -1;
        if (is_float) {
            int64_t acc = 0;
            for (int32_t i = 0; i < n; ++i) {
                double lj = 0.0;
                for (int32_t t = rp[i]; t < rp[i + 1]; ++t)
                    lj = sa[ci[t]] > 0 ? lj + (double)wf[t] : lj - (double)wf[t];
                double v = 0.5 * lj + (hf ? (double)hf[i] : 0.0);
                double e = sa[i] > 0 ? -v : v;
                acc += fix36d(e);
            }
            E[a] = acc;
        } else {
            int64_t s1 = 0, s2 = 0;
            for (int32_t i = 0; i < n; ++i) {
                int64_t lj = 0;
                for (int32_t t = rp[i]; t < rp[i + 1]; ++t) {
                    int64_t w = P->weight_type == OR_W_I8 ? w8[t] : w32[t];
                    lj += sa[ci[t]] > 0 ? w : -w;
                }
                s1 += sa[i] > 0 ? lj : -lj;
                if (hi) s2 += sa[i] > 0 ? hi[i] : -hi[i];
            }
            E[a] = -(s1 / 2) - s2;
        }
        rung_of[a] = a;
        rep_at[a] = a;
    }
    if (R > 1) { label[rep_at[0]] = 1; label[rep_at[R - 1]] = -1; }
    for (int32_t r = 0; r + 1 < R; ++r) {
        if (res->swap_tried) res->swap_tried[r] = 0;
        if (res->swap_acc) res->swap_acc[r] = 0;
    }
    for (int32_t r = 0; r < R; ++r) {
        if (res->round_trips) res->round_trips[r] = 0;
        if (res->rung_up) res->rung_up[r] = 0;
        if (res->rung_down) res->rung_down[r] = 0;
    }

    /* best over rungs in order, strict < (round 0 = the initial states) */
    int have_best = 0;
    int64_t bestu = 0;
    int32_t best_round = 0, best_rung = 0, best_slot = 0;
#define OR_SCAN_BEST(RND)                                                              \
    for (int32_t r = 0; r < R; ++r) {                                                  \
        int32_t a = rep_at[r];                                                         \
        if (!have_best || E[a] < bestu) {                                              \
            have_best = 1; bestu = E[a]; best_round = (RND); best_rung = r; best_slot = a; \
            memcpy(res->best_s, s + (size_t)a * n, (size_t)n);                         \
        }                                                                              \
    }
    OR_SCAN_BEST(0)

#ifdef _OPENMP
    if (prm->n_threads > 0) omp_set_num_threads(prm->n_threads);
#endif
    double t0 = now_s();
    for (int32_t rnd = 0; rnd < prm->n_rounds; ++rnd) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t a = 0; a < R; ++a) {
            int8_t *sa = s + (size_t)a * n;
            const int32_t rung = rung_of[a];
            const float beta_f = (float)prm->beta[rung];
            const uint32_t *thr_r = thr ? thr + (size_t)rung * qmax : NULL;
            int64_t dE_acc = 0;
            for (int32_t j = 0; j < S; ++j) {
                const uint32_t t = (uint32_t)(rnd * S + j);
                for (int32_t c = 0; c < P->n_colors; ++c) {
                    uint32_t word[4] = {0, 0, 0, 0};
                    int32_t have_quad = -1;
                    for (int32_t p = P->color_ptr[c]; p < P->color_ptr[c + 1]; ++p) {
                        if ((p >> 2) != have_quad) {
                            uint32_t ctr[4] = {(uint32_t)(p >> 2), (uint32_t)a, t, 2u};
                            uint32_t key[2] = {(uint32_t)prm->seed, (uint32_t)(prm->seed >> 32)};
                            or_philox(ctr, key, word);
                            have_quad = p >> 2;
                        }
                        const uint32_t u = word[p & 3];
                        const int32_t i = P->color_spins[p];
                        const int8_t si = sa[i];
                        if (is_float) {
                            /* the field of float32 couplings summed in double (exact for
                             * magnitudes within 2^29 of one another), the decision on its
                             * float rounding, the energy change from the double */
                            double l = hf ? (double)hf[i] : 0.0;
                            for (int32_t tt = rp[i]; tt < rp[i + 1]; ++tt)
                                l = sa[ci[tt]] > 0 ? l + (double)wf[tt] : l - (double)wf[tt];
                            double q = si > 0 ? l : -l;
                            if (metro_float(beta_f, (float)q, u)) {
                                sa[i] = (int8_t)-si;
                                dE_acc += fix36d(2.0 * q);
                            }
                        } else {
                            int64_t l = hi ? hi[i] : 0;
                            for (int32_t tt = rp[i]; tt < rp[i + 1]; ++tt) {
                                int64_t w = P->weight_type == OR_W_I8 ? w8[tt] : w32[tt];
                                l += sa[ci[tt]] > 0 ? w : -w;
                            }
                            int64_t q = si > 0 ? l : -l;
                            int acc = q <= 0 || u < thr_r[q - 1];
                            if (acc) {
                                sa[i] = (int8_t)-si;
                                dE_acc += 2 * q;
                            }
                        }
                    }
                }
            }
            E[a] += dE_acc;
        }
        OR_SCAN_BEST(rnd + 1)
        if (res->best_history) res->best_history[rnd] = (double)bestu * unit;
        if (res->rung_trace)
            for (int32_t r = 0; r < R; ++r)
                res->rung_trace[(size_t)rnd * R + r] = (double)E[rep_at[r]] * unit;
        /* exchange: pairs (r, r+1), r = rnd%2, rnd%2+2, ... (parallel_tempering.py:104-113) */
        for (int32_t r = rnd % 2; r + 1 < R; r += 2) {
            int32_t a = rep_at[r], b = rep_at[r + 1];
            uint32_t u = philox_word((uint32_t)(r >> 2), 0u, (uint32_t)rnd, 3u, prm->seed, r);
            int acc = exch_accept(prm->beta[r], prm->beta[r + 1], (double)E[a] * unit,
                                  (double)E[b] * unit, u);
            if (res->swap_tried) res->swap_tried[r]++;
            if (acc) {
                if (res->swap_acc) res->swap_acc[r]++;
                rep_at[r] = b; rep_at[r + 1] = a;
                rung_of[a] = r + 1; rung_of[b] = r;
            }
        }
        if (R > 1) {
            for (int32_t a = 0; a < R; ++a) {
                if (rung_of[a] == 0) {
                    if (label[a] == -1 && res->round_trips) res->round_trips[a]++;
                    label[a] = 1;
                } else if (rung_of[a] == R - 1) {
                    label[a] = -1;
                }
                if (label[a] == 1 && res->rung_up) res->rung_up[rung_of[a]]++;
                if (label[a] == -1 && res->rung_down) res->rung_down[rung_of[a]]++;
            }
        }
    }
    res->seconds = now_s() - t0;
#undef OR_SCAN_BEST
    res->best_units = bestu;
    res->best_E = (double)bestu * unit;
    res->best_round = best_round;
    res->best_rung = best_rung;
    res->best_slot = best_slot;
    for (int32_t a = 0; a < R; ++a) {
        if (res->rung_of) res->rung_of[a] = rung_of[a];
        if (res->final_units) res->final_units[a] = E[a];
    }
    if (res->final_spins) memcpy(res->final_spins, s, (size_t)R * n);
    free(thr); free(s); free(E); free(rung_of); free(rep_at); free(label);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Exact enumeration for n <= 24 (drift/solvers/exact.py:20-49): Gray-code walk with maintained
 * local fields.  Configuration index idx has bit b <-> spin (1 - 2b); the minimum is the
 * lexicographically first (E, idx), matching numpy argmin over the natural order.  Also returns
 * the exact Boltzmann averages <E> and <E^2> at each beta.
 * ---------------------------------------------------------------------------------------- */
int32_t or_exact(int32_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *w,
                 const double *h, int32_t n_beta, const double *betas, double *e_min,
                 int64_t *idx_min, double *mean_E, double *mean_E2) {
    if (n < 1 || n > 26) return -1;
    double *l = (double *)malloc(sizeof(double) * (size_t)n);
    int8_t *s = (int8_t *)malloc((size_t)n);
    double emin = 0.0;
    int64_t imin = 0;
    for (int pass = 0; pass < 2; ++pass) {
        double *Z = (double *)calloc((size_t)(n_beta > 0 ? n_beta : 1), sizeof(double));
        double *ZE = (double *)calloc((size_t)(n_beta > 0 ? n_beta : 1), sizeof(double));
        double *ZE2 = (double *)calloc((size_t)(n_beta > 0 ? n_beta : 1), sizeof(double));
        for (int32_t i = 0; i < n; ++i) s[i] = 1;
        double E = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            double lj = 0.0;
            for (int32_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t) lj += w[t] * s[col_idx[t]];
            l[i] = lj + (h ? h[i] : 0.0);
            E += -0.5 * s[i] * lj - (h ? h[i] : 0.0) * s[i];
        }
        int64_t idx = 0;
        const int64_t total = (int64_t)1 << n;
        for (int64_t k = 0; k < total; ++k) {
            if (k > 0) {
                int b = __builtin_ctzll((unsigned long long)k);
                E += 2.0 * s[b] * l[b];
                s[b] = (int8_t)-s[b];
                idx ^= (int64_t)1 << b;
                for (int32_t t = row_ptr[b]; t < row_ptr[b + 1]; ++t)
                    l[col_idx[t]] += 2.0 * w[t] * s[b];
                if ((k & 4095) == 0) {   /* re-anchor to bound floating drift */
                    double Ex = 0.0;
                    for (int32_t i = 0; i < n; ++i) {
                        double lj = 0.0;
                        for (int32_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t)
                            lj += w[t] * s[col_idx[t]];
                        l[i] = lj + (h ? h[i] : 0.0);
                        Ex += -0.5 * s[i] * lj - (h ? h[i] : 0.0) * s[i];
                    }
                    E = Ex;
                }
            }
            if (pass == 0) {
                if (k == 0 || E < emin || (E == emin && idx < imin)) { emin = E; imin = idx; }
            } else {
                double x = E - emin;
                for (int32_t bi = 0; bi < n_beta; ++bi) {
                    double wgt = exp(-betas[bi] * x);
                    Z[bi] += wgt;
                    ZE[bi] += wgt * E;
                    ZE2[bi] += wgt * E * E;
                }
            }
        }
        if (pass == 1)
            for (int32_t bi = 0; bi < n_beta; ++bi) {
                if (mean_E) mean_E[bi] = ZE[bi] / Z[bi];
                if (mean_E2) mean_E2[bi] = ZE2[bi] / Z[bi];
            }
        free(Z); free(ZE); free(ZE2);
        if (n_beta <= 0) break;
    }
    *e_min = emin;
    *idx_min = imin;
    free(l); free(s);
    return 0;
}
