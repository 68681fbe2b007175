// pt_device.h -- device helpers shared by the engines: best-rung scan, exchange, round-trip
// labels and the initial spin of the stream contract.  Internal to libdriftpt.so.
#pragma once
#include <climits>

#include "pt_engine.h"
#include "pt_rules.h"

#define PT_WAVE 64

namespace pt_dev {

__device__ __forceinline__ void lex_min(long long& e, int& r, long long e2, int r2) {
    if (e2 < e || (e2 == e && r2 < r)) {
        e = e2;
        r = r2;
    }
}

// Block-wide lexicographic minimum of (E[rep[r]], r) over rungs r < R.  Every thread calls it and
// receives the result; this is the "scan rungs in order, strict <" rule of the reference's best
// tracking (drift/solvers/parallel_tempering.py:99-102) evaluated in parallel.
__device__ inline void block_best_rung(const int64_t* E, const int32_t* rep, int R, long long* scrE,
                                int* scrR, int64_t& outE, int32_t& outR) {
    long long e = LLONG_MAX;
    int r = INT_MAX;
    for (int x = threadIdx.x; x < R; x += blockDim.x) lex_min(e, r, (long long)E[rep[x]], x);
#pragma unroll
    for (int m = PT_WAVE / 2; m > 0; m >>= 1) {
        const long long e2 = __shfl_xor(e, m);
        const int r2 = __shfl_xor(r, m);
        lex_min(e, r, e2, r2);
    }
    const int wid = threadIdx.x / PT_WAVE, nw = (blockDim.x + PT_WAVE - 1) / PT_WAVE;
    if ((threadIdx.x & (PT_WAVE - 1)) == 0) {
        scrE[wid] = e;
        scrR[wid] = r;
    }
    __syncthreads();
    e = scrE[0];
    r = scrR[0];
    for (int w = 1; w < nw; ++w) lex_min(e, r, scrE[w], scrR[w]);
    outE = e;
    outR = r;
    __syncthreads();
}

__device__ __forceinline__ int8_t init_spin(const PtInstance& I, int p, int a) {
    const int i = I.orig[p];
    const pt_u4 r = pt_philox((uint32_t)(i >> 2), (uint32_t)a, 0u, PT_TAG_INIT, I.k0, I.k1);
    return (pt_word(r, i & 3) & 1u) ? (int8_t)1 : (int8_t)-1;
}

// Exchange of one round for the pairs (r, r+1), r = rnd%2 + 2m (parallel_tempering.py:104-113),
// with the permutation rung<->slot instead of moving configurations.  Every pair is disjoint, so
// each thread owns one pair.  Caller syncs before and after; null counters are not updated.
__device__ __forceinline__ void exchange_pairs(const PtInstance& I, int rnd, const int64_t* E,
                                               int32_t* rep, int32_t* rung, int64_t* tried,
                                               int64_t* accd) {
    const int R = I.R;
    for (int m = threadIdx.x;; m += blockDim.x) {
        const int r = (rnd & 1) + 2 * m;
        if (r + 1 >= R) break;
        const int a = rep[r], b = rep[r + 1];
        const pt_u4 rv = pt_philox((uint32_t)(r >> 2), 0u, (uint32_t)rnd, PT_TAG_EXCH, I.k0, I.k1);
        const bool ok = pt_accept_exchange(I.beta[r], I.beta[r + 1], (double)E[a] * I.unit,
                                           (double)E[b] * I.unit, pt_word(rv, r & 3));
        if (tried) tried[r] += 1;
        if (ok) {
            if (accd) accd[r] += 1;
            rep[r] = b;
            rep[r + 1] = a;
            rung[a] = r + 1;
            rung[b] = r;
        }
    }
}

// Round-trip labels after an exchange: a slot reaching the coldest rung with a "hot" label
// completes a round trip.  rung is a permutation, so per-rung counters have one writer each.
__device__ __forceinline__ void update_labels(int R, const int32_t* rung, int32_t* label,
                                              int32_t* trips32, int64_t* trips64, int64_t* up,
                                              int64_t* down) {
    if (R < 2) return;
    for (int a = threadIdx.x; a < R; a += blockDim.x) {
        const int g = rung[a];
        int lb = label[a];
        if (g == 0) {
            if (lb == -1) {
                if (trips32) trips32[a] += 1;
                if (trips64) trips64[a] += 1;
            }
            lb = 1;
        } else if (g == R - 1) {
            lb = -1;
        }
        label[a] = lb;
        if (lb == 1) up[g] += 1;
        if (lb == -1) down[g] += 1;
    }
}

}  // namespace pt_dev
