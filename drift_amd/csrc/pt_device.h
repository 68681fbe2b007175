// pt_device.h -- device helpers shared by the engines: best-rung scan, exchange, round-trip
// labels, the initial spin of the stream contract and the CSR site update.  Internal to
// libdriftpt.so.
#pragma once
#include <climits>
#include <type_traits>

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

// Counter increment without waiting for the old value (each counter has one writer per round;
// the atomic only spares the dependent load-add-store round trip to memory).
__device__ __forceinline__ void bump(int64_t* p) { atomicAdd((unsigned long long*)p, 1ull); }

__device__ __forceinline__ int8_t init_spin(const PtInstance& I, int p, int a) {
    const int i = I.orig[p];
    const pt_u4 r = pt_philox((uint32_t)(i >> 2), (uint32_t)a, 0u, PT_TAG_INIT, I.k0, I.k1);
    return (pt_word(r, i & 3) & 1u) ? (int8_t)1 : (int8_t)-1;
}

// Exchange of one round for the pairs (r, r+1), r = rnd%2 + 2m (parallel_tempering.py:104-113),
// with the permutation rung<->slot instead of moving configurations.  Every pair is disjoint, so
// each thread owns one pair.  Caller syncs before and after; null counters are not updated.
// beta may be a copy of the instance's ladder in LDS.
__device__ __forceinline__ void exchange_pairs(const PtInstance& I, const double* beta, int rnd,
                                               const int64_t* E, int32_t* rep, int32_t* rung,
                                               int64_t* tried, int64_t* accd) {
    const int R = I.R;
    for (int m = threadIdx.x;; m += blockDim.x) {
        const int r = (rnd & 1) + 2 * m;
        if (r + 1 >= R) break;
        const int a = rep[r], b = rep[r + 1];
        const pt_u4 rv = pt_philox((uint32_t)(r >> 2), 0u, (uint32_t)rnd, PT_TAG_EXCH, I.k0, I.k1);
        const bool ok = pt_accept_exchange(beta[r], beta[r + 1], (double)E[a] * I.unit,
                                           (double)E[b] * I.unit, pt_word(rv, r & 3));
        if (tried) bump(tried + r);
        if (ok) {
            if (accd) bump(accd + r);
            rep[r] = b;
            rep[r + 1] = a;
            rung[a] = r + 1;
            rung[b] = r;
        }
    }
}

__device__ __forceinline__ void exchange_pairs(const PtInstance& I, int rnd, const int64_t* E,
                                               int32_t* rep, int32_t* rung, int64_t* tried,
                                               int64_t* accd) {
    exchange_pairs(I, I.beta, rnd, E, rep, rung, tried, accd);
}

// Round-trip labels after an exchange: a slot reaching the coldest rung with a "hot" label
// completes a round trip.  rung is a permutation, so per-rung counters have one writer each.
// Null counters are not updated.
__device__ __forceinline__ void update_labels(int R, const int32_t* rung, int32_t* label,
                                              int32_t* trips32, int64_t* trips64, int64_t* up,
                                              int64_t* down) {
    if (R < 2) return;
    for (int a = threadIdx.x; a < R; a += blockDim.x) {
        const int g = rung[a];
        int lb = label[a];
        if (g == 0) {
            if (lb == -1) {
                if (trips32) atomicAdd(trips32 + a, 1);
                if (trips64) bump(trips64 + a);
            }
            lb = 1;
        } else if (g == R - 1) {
            lb = -1;
        }
        label[a] = lb;
        if (up && lb == 1) bump(up + g);
        if (down && lb == -1) bump(down + g);
    }
}

// Edge t of the CSR (row order): neighbour position and weight.
template <int WM>
__device__ __forceinline__ int edge_int(const void* __restrict__ e, int t, int32_t& wt) {
    if constexpr (WM == PT_WM_I8) {
        const uint32_t v = ((const uint32_t*)e)[t];
        wt = (int32_t)(int8_t)(v & 0xffu);
        return (int)(v >> 8);
    } else {
        const uint2 v = ((const uint2*)e)[t];
        wt = (int32_t)v.y;
        return (int)v.x;
    }
}

__device__ __forceinline__ int edge_float(const void* __restrict__ e, int t, float& wt) {
    const uint2 v = ((const uint2*)e)[t];
    wt = __uint_as_float(v.y);
    return (int)v.x;
}

// Sign bit of a +-1 spin as a float's sign: XORed into a coupling's bits it gives J * s exactly.
__device__ __forceinline__ uint32_t pt_sign_bit(int8_t s) { return (uint32_t)(int32_t)s & 0x80000000u; }

// exp(x) from the hardware base-2 exponential (v_exp_f32), for x <= 0.
__device__ __forceinline__ float pt_exp_hw(float x) {
    return __builtin_amdgcn_exp2f(x * 0x1.715476p+0f);
}

// pt_accept_float with the same result, decided from the hardware exponential where that cannot
// differ.  x = (-2 beta) q is pt_accept_float's -(beta (2 q)) (one rounding of the same product;
// the host keeps beta below 2^100 so 2 beta is exact).  With pa = pt_exp_hw(x) and
// p = pt_expf_det(x), |p - pa| <= 2^-13 pa for every float x in (-87, 0]
// (drift_probe_exp_margin checks this on the device for all of them).  Then for q > 0
//   u < floor(p 2^32)  whenever  (float)u < fma(pa, 2^32 (1 - 2^-12), -512),  and
//   u >= floor(p 2^32) whenever  (float)u > fma(pa, 2^32 (1 + 2^-12), 512):
// the relative margin 2^-12 exceeds the 2^-13 gap plus the bound's rounding (2^-24), and 512
// exceeds the rounding of (float)u (at most 128) plus the floor (1) plus 128.  For x <= -87,
// p = 0 and the first test fails while the second holds for u > 512; p >= 1 (threshold 2^32 - 1)
// only where pa > 1 - 2^-13, so the second test never holds there.  For q <= 0 (always accepted)
// pa >= 1 - 2^-13 as well, or +inf, so the second test fails and the first holds outside the band.
// Only u in the band between the bounds (about 2^-11 of them) takes the exact polynomial.
__device__ __forceinline__ bool pt_accept_float_dev(float beta, float q, uint32_t u) {
    const float pa = pt_exp_hw((-2.0f * beta) * q);
    const float uf = (float)u;
    const bool acc = uf < __builtin_fmaf(pa, 0x1.ffep+31f, -512.0f);
    const bool rej = uf > __builtin_fmaf(pa, 0x1.001p+32f, 512.0f);
    if (__builtin_expect(acc == rej, 0)) return pt_accept_float(beta, q, u);
    return acc;
}

// pt_fix36d(2 v): 2^37 v is exact, so one fused multiply-add gives the same rounded sum.
__device__ __forceinline__ int64_t pt_fix37d(double v) {
    union {
        double d;
        int64_t i;
    } c;
    c.d = __builtin_fma(v, 137438953472.0, 6755399441055744.0);
    return c.i - (int64_t)0x4338000000000000ll;
}

// Spin access.  COH accesses are agent-scope relaxed atomics (loads and stores coherent across
// the chip without cache maintenance): the coop engine's workgroups read spins other workgroups
// wrote in the same launch.  Other engines use plain accesses.
template <bool COH>
__device__ __forceinline__ int8_t ld_spin(const int8_t* p) {
    if constexpr (COH)
        return __hip_atomic_load(const_cast<int8_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}

template <bool COH>
__device__ __forceinline__ void st_spin(int8_t* p, int8_t v) {
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// Local field of position p in slot a: the sum over its row of w_t * s_col(t), in row order.
// Rows are gathered in chunks of G edges without branches: a slot past the end of the row reads
// edge 0 and position 0 (valid addresses) and is dropped by a select, so every load of a chunk
// issues back to back and the chunk waits on memory once.  Spins in HBM use G = 8; spins in LDS
// use G = 1 (the plain loop), where the extra lanes only cost issue slots.
template <int WM, int G, bool COH = false>
__device__ __forceinline__ int32_t field_int(int a, int Rp, const int8_t* sp, int t0, int t1,
                                             const void* __restrict__ edge, int32_t l) {
    for (int t = t0; t < t1; t += G) {
        int c[G];
        int32_t w[G];
        int8_t v[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const bool ok = t + u < t1;
            c[u] = edge_int<WM>(edge, ok ? t + u : 0, w[u]);
            c[u] = ok ? c[u] : 0;
        }
#pragma unroll
        for (int u = 0; u < G; ++u) v[u] = ld_spin<COH>(sp + (size_t)c[u] * Rp + a);
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int32_t term = v[u] > 0 ? w[u] : -w[u];
            l += t + u < t1 ? term : 0;
        }
    }
    return l;
}

template <int G, bool COH = false>
__device__ __forceinline__ double field_float(int a, int Rp, const int8_t* sp, int t0, int t1,
                                              const void* __restrict__ edge, double l) {
    for (int t = t0; t < t1; t += G) {
        int c[G];
        float w[G];
        int8_t v[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const bool ok = t + u < t1;
            c[u] = edge_float(edge, ok ? t + u : 0, w[u]);
            c[u] = ok ? c[u] : 0;
        }
#pragma unroll
        for (int u = 0; u < G; ++u) v[u] = ld_spin<COH>(sp + (size_t)c[u] * Rp + a);
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const double nl = l + (double)__uint_as_float(__float_as_uint(w[u]) ^ pt_sign_bit(v[u]));
            l = t + u < t1 ? nl : l;
        }
    }
    return l;
}

// The attempts of one Philox quad in one slot: positions 4q .. 4q+3 that lie in colour range
// [cs, ce), with word k of rv for position 4q+k.  The first G edge records of every row are also
// stored row-major in ell, with the neighbour as its LDS offset (position * Rp) and zero records
// padding short rows (offset 0, weight 0), so the quad's first chunks are 4G contiguous records
// read with wide loads, no per-edge bounds test and no address multiply; h is padded to a multiple
// of 4 positions and read as one 16-byte load.  Same-colour positions share no edge, so all four
// rows are gathered before any spin of the quad changes: the records and the neighbour spins each
// wait on memory once per quad.  Only when some row is longer than G (rest) are the row pointers
// read and the longer rows finished from the CSR in G-edge chunks.  Each field is summed from h in
// row order (float couplings in double), one rounding per term (l +- w), as field_float /
// field_int sum it and as the oracle does; a padding record adds +0 or -0, which changes at most the sign of a zero field, and a zero
// field is accepted with zero energy change either way, so the decisions are the same.  Positions
// of the quad outside [cs, ce) are gathered too (valid addresses) and then left alone.
// Loads through global-address-space pointers wait on their own counter, while flat loads (what a
// pointer read from memory compiles to) also hold up the waits for LDS reads issued after them.
#define PT_GLOBAL __attribute__((address_space(1)))
typedef uint32_t pt_v4u __attribute__((ext_vector_type(4)));

template <int WM, int G>
__device__ __forceinline__ int64_t quad_update(int q, int cs, int ce, int n, int a, int Rp,
                                               int8_t* sp, const int32_t* __restrict__ rp,
                                               const void* __restrict__ edge,
                                               const void* __restrict__ ell, const void* h,
                                               bool rest, const pt_u4& rv,
                                               const uint32_t* thr_row, int TQ, float beta_f) {
    using LT = typename std::conditional<WM == PT_WM_F32, double, int32_t>::type;
    constexpr int EW = WM == PT_WM_I8 ? 1 : 2;   // 32-bit words per edge record
    constexpr int NW = 4 * G * EW;                // words of the quad's first chunks
    uint32_t ev[NW];
    {
        const PT_GLOBAL pt_v4u* src = (const PT_GLOBAL pt_v4u*)ell + (size_t)q * (NW / 4);
#pragma unroll
        for (int i = 0; i < NW / 4; ++i) {
            const pt_v4u x = src[i];
            ev[4 * i] = x.x;
            ev[4 * i + 1] = x.y;
            ev[4 * i + 2] = x.z;
            ev[4 * i + 3] = x.w;
        }
    }
    // int8 couplings, rows longer than the chunk (sparse graphs of mean degree up to 12): their
    // row pointers, issued with the chunk's records so the remainder's first pass does not wait on
    // another round trip for them (on a batch of such graphs, 65.3 -> 63.0 s); the wider records
    // keep the load after the chunk, where its registers are free
    int rb[5] = {0, 0, 0, 0, 0};
    if (WM == PT_WM_I8 && rest) {
#pragma unroll
        for (int k = 0; k < 5; ++k) rb[k] = ((const PT_GLOBAL int32_t*)rp)[min(4 * q + k, n)];
    }
    LT l[4] = {(LT)0, (LT)0, (LT)0, (LT)0};
    if (h) {
        const pt_v4u hv = ((const PT_GLOBAL pt_v4u*)h)[q];
        if constexpr (WM == PT_WM_F32) {
            l[0] = (double)__uint_as_float(hv.x); l[1] = (double)__uint_as_float(hv.y);
            l[2] = (double)__uint_as_float(hv.z); l[3] = (double)__uint_as_float(hv.w);
        } else {
            l[0] = (int32_t)hv.x; l[1] = (int32_t)hv.y; l[2] = (int32_t)hv.z; l[3] = (int32_t)hv.w;
        }
    }
    // positions past n (the last quad) read LDS words that follow the spins and are not stored
    int8_t* own = sp + (uint32_t)(4 * q * Rp + a);
    int8_t si[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) si[k] = own[k * Rp];
    {
        int8_t v[4 * G];
#pragma unroll
        for (int e = 0; e < 4 * G; ++e)
            v[e] = sp[(WM == PT_WM_I8 ? (ev[e] >> 8) : ev[EW * e]) + a];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int e = k * G + u;
                // the spin is +-1, so w * s is w with the spin's sign bit XORed in: exact, and
                // l + w*s rounds once, as the select of l + w and l - w does
                if constexpr (WM == PT_WM_F32) {
                    l[k] += (double)__uint_as_float(ev[EW * e + 1] ^ pt_sign_bit(v[e]));
                } else if constexpr (WM == PT_WM_I8) {
                    l[k] += __mul24((int32_t)(int8_t)(ev[e] & 0xffu), (int32_t)v[e]);
                } else {
                    const int32_t w = (int32_t)ev[EW * e + 1];
                    l[k] += v[e] > 0 ? w : -w;
                }
            }
    }
    bool on[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) on[k] = 4 * q + k >= cs && 4 * q + k < ce;
    if (rest) {
        // the rows' remainders in chunks, the four rows' chunks together: each pass issues every
        // row's next records at once and waits on memory once, then on LDS once, so a quad pays
        // one round trip per pass instead of one per row and chunk.  Each row still sums its
        // terms in row order, one rounding each, as field_float / field_int do.
        if constexpr (WM != PT_WM_I8) {
#pragma unroll
            for (int k = 0; k < 5; ++k) rb[k] = ((const PT_GLOBAL int32_t*)rp)[min(4 * q + k, n)];
        }
        int t[4], e[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            t[k] = rb[k] + G;
            e[k] = on[k] ? rb[k + 1] : 0;
        }
        // RC records per row and pass keep the pass within the register budget (4 RC records,
        // their positions, weights and spins live at once; 3 spilled on the split kernel, and 3
        // or 4 for int8 records, which did not spill, ran a batch of sparse ER graphs 4% and 7%
        // slower)
        constexpr int RC = 2;
        while ((t[0] < e[0]) | (t[1] < e[1]) | (t[2] < e[2]) | (t[3] < e[3])) {
            int c[4 * RC];
            int8_t v[4 * RC];
            using WT = typename std::conditional<WM == PT_WM_F32, float, int32_t>::type;
            WT w[4 * RC];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int u = 0; u < RC; ++u) {
                    const bool ok = t[k] + u < e[k];
                    if constexpr (WM == PT_WM_F32) c[k * RC + u] = edge_float(edge, ok ? t[k] + u : 0, w[k * RC + u]);
                    else c[k * RC + u] = edge_int<WM>(edge, ok ? t[k] + u : 0, w[k * RC + u]);
                    c[k * RC + u] = ok ? c[k * RC + u] : 0;
                }
            // all records, then all spins: without the barriers the scheduler pairs each record
            // with its spin and waits in between
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int x = 0; x < 4 * RC; ++x) v[x] = sp[(size_t)c[x] * Rp + a];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int u = 0; u < RC; ++u) {
                    const bool ok = t[k] + u < e[k];
                    if constexpr (WM == PT_WM_F32) {
                        const double nl = l[k] + (double)__uint_as_float(
                            __float_as_uint(w[k * RC + u]) ^ pt_sign_bit(v[k * RC + u]));
                        l[k] = ok ? nl : l[k];
                    } else {
                        const int32_t term = v[k * RC + u] > 0 ? w[k * RC + u] : -w[k * RC + u];
                        l[k] += ok ? term : 0;
                    }
                }
                t[k] += RC;
            }
        }
    }
    int64_t d = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!on[k]) continue;
        if constexpr (WM == PT_WM_F32) {
            const double qd = si[k] > 0 ? l[k] : -l[k];
            if (pt_accept_float_dev(beta_f, (float)qd, pt_word(rv, k))) {
                own[k * Rp] = (int8_t)-si[k];
                d += pt_fix37d(qd);
            }
        } else {
            const int32_t qi = si[k] > 0 ? l[k] : -l[k];
            if (qi <= 0 || (qi <= TQ && pt_word(rv, k) < thr_row[qi - 1])) {
                own[k * Rp] = (int8_t)-si[k];
                d += 2 * (int64_t)qi;
            }
        }
    }
    return d;
}

// The attempts of one Philox quad in one slot with the instance's couplings in LDS (split engine,
// int8 couplings with |w| <= 7 and n <= 4096): row pointers rpl (n + 1), 16-bit records recl
// (position << 4 | weight & 15, in row order), field hl (n, or null).  Every gather is an LDS
// read: the four rows' records are taken four per row and pass, all rows together, then their
// spins.  The decisions are quad_update's on the integer path (an integer field is exact in any
// order of summation).
__device__ __forceinline__ int64_t quad_update_cl(int q, int cs, int ce, int a, int Rp, int8_t* sp,
                                                  const int32_t* rpl, const uint16_t* recl,
                                                  const int32_t* hl, const pt_u4& rv,
                                                  const uint32_t* thr_row, int TQ) {
    int8_t* own = sp + (uint32_t)(4 * q * Rp + a);
    int8_t si[4];
    int32_t l[4];
    int t[4], e[4];
    bool on[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = 4 * q + k;
        on[k] = p >= cs && p < ce;
        si[k] = own[k * Rp];
        t[k] = on[k] ? rpl[p] : 0;
        e[k] = on[k] ? rpl[p + 1] : 0;
        l[k] = (on[k] && hl) ? hl[p] : 0;
    }
    while ((t[0] < e[0]) | (t[1] < e[1]) | (t[2] < e[2]) | (t[3] < e[3])) {
        uint32_t r[16];
        int8_t v[16];
        // past a row's end a lane reads record 0 (it exists whenever some row is not done) and
        // drops its term
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int u = 0; u < 4; ++u) r[4 * k + u] = recl[t[k] + u < e[k] ? t[k] + u : 0];
        // the scheduler otherwise pairs each record load with its spin load and waits between
        // them (16 LDS round trips per pass instead of 2)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = sp[(r[x] >> 4) * (uint32_t)Rp + a];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t w = (int32_t)(r[4 * k + u] << 28) >> 28;
                const int32_t term = v[4 * k + u] > 0 ? w : -w;
                l[k] += t[k] + u < e[k] ? term : 0;
            }
            t[k] += 4;
        }
    }
    int64_t d = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!on[k]) continue;
        const int32_t qi = si[k] > 0 ? l[k] : -l[k];
        if (qi <= 0 || (qi <= TQ && pt_word(rv, k) < thr_row[qi - 1])) {
            own[k * Rp] = (int8_t)-si[k];
            d += 2 * (int64_t)qi;
        }
    }
    return d;
}

// Site update of position p in slot a with spins in HBM [position][slot]: the first GW edge
// records of the row from the padded row-major copy (neighbour as its position), the rest of a
// longer row (when some row is, rest) from the CSR in GW-edge chunks.  The records do not depend
// on the row pointers, so the chain is record -> neighbour spin.  Same decision and field order
// as quad_update; a padding record adds +-0 (see there).
template <int WM, int GW>
__device__ __forceinline__ int64_t site_update_ell(int p, int a, int Rp, int8_t* sp,
                                                   const int32_t* __restrict__ rp,
                                                   const void* __restrict__ edge,
                                                   const void* __restrict__ ell, bool rest,
                                                   const void* h, uint32_t u,
                                                   const uint32_t* thr_row, int TQ, float beta_f) {
    using LT = typename std::conditional<WM == PT_WM_F32, double, int32_t>::type;
    constexpr int EW = WM == PT_WM_I8 ? 1 : 2;
    constexpr int NW = GW * EW;
    uint32_t ev[NW];
    {
        const PT_GLOBAL pt_v4u* src = (const PT_GLOBAL pt_v4u*)ell + (size_t)p * (NW / 4);
#pragma unroll
        for (int i = 0; i < NW / 4; ++i) {
            const pt_v4u x = src[i];
            ev[4 * i] = x.x;
            ev[4 * i + 1] = x.y;
            ev[4 * i + 2] = x.z;
            ev[4 * i + 3] = x.w;
        }
    }
    int t0 = 0, t1 = 0;
    if (rest) {
        t0 = rp[p];
        t1 = rp[p + 1];
    }
    int8_t* own = sp + (size_t)p * Rp + a;
    const int8_t si = *own;
    LT l = (LT)0;
    if (h) {
        if constexpr (WM == PT_WM_F32) l = (double)((const float*)h)[p];
        else l = ((const int32_t*)h)[p];
    }
    int8_t v[GW];
#pragma unroll
    for (int e = 0; e < GW; ++e)
        v[e] = sp[(size_t)(WM == PT_WM_I8 ? (ev[e] >> 8) : ev[EW * e]) * Rp + a];
#pragma unroll
    for (int e = 0; e < GW; ++e) {
        if constexpr (WM == PT_WM_F32) {
            l += (double)__uint_as_float(ev[EW * e + 1] ^ pt_sign_bit(v[e]));
        } else if constexpr (WM == PT_WM_I8) {
            l += __mul24((int32_t)(int8_t)(ev[e] & 0xffu), (int32_t)v[e]);
        } else {
            const int32_t w = (int32_t)ev[EW * e + 1];
            l += v[e] > 0 ? w : -w;
        }
    }
    if (t0 + GW < t1) {
        if constexpr (WM == PT_WM_F32) l = field_float<GW>(a, Rp, sp, t0 + GW, t1, edge, l);
        else l = field_int<WM, GW>(a, Rp, sp, t0 + GW, t1, edge, l);
    }
    if constexpr (WM == PT_WM_F32) {
        const double q = si > 0 ? l : -l;
        if (pt_accept_float_dev(beta_f, (float)q, u)) {
            *own = (int8_t)-si;
            return pt_fix37d(q);
        }
        return 0;
    } else {
        const int32_t q = si > 0 ? l : -l;
        if (q <= 0 || (q <= TQ && u < thr_row[q - 1])) {
            *own = (int8_t)-si;
            return 2 * (int64_t)q;
        }
        return 0;
    }
}

// Energy contribution of position p in slot a at initialisation.  Integer modes return
// s_p (sum_j J s_j + 2 h_p) (E = -total/2); float returns fix36(-s_p (l_J/2 + h_p)), summed in
// double.
template <int WM, int G = 1>
__device__ __forceinline__ int64_t site_energy(int p, int a, int Rp, const int8_t* sp,
                                               const int32_t* __restrict__ rp,
                                               const void* __restrict__ edge, const void* h) {
    const int t0 = rp[p], t1 = rp[p + 1];
    const int8_t si = sp[(size_t)p * Rp + a];
    if constexpr (WM == PT_WM_F32) {
        const double lj = field_float<G>(a, Rp, sp, t0, t1, edge, 0.0);
        const double v = 0.5 * lj + (h ? (double)((const float*)h)[p] : 0.0);
        return pt_fix36d(si > 0 ? -v : v);
    } else {
        // |sum_j J s_j| < 2^31 is checked on the host (maximum field), so int32 is exact
        const int64_t lj = field_int<WM, G>(a, Rp, sp, t0, t1, edge, 0);
        const int64_t hh = h ? (int64_t)((const int32_t*)h)[p] : 0;
        const int64_t v = lj + 2 * hh;
        return si > 0 ? v : -v;
    }
}

// Persistent engines (coop, pslab): every access one workgroup makes to data another workgroup
// wrote in the same launch (spins, energy sums, barrier flags) is an agent-scope relaxed atomic,
// which gfx950 compiles to sc1 loads and stores: coherent across the chip without cache
// maintenance.
__device__ __forceinline__ uint32_t ld_u32(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t ld_coh(const int64_t* p) {
    return __hip_atomic_load(const_cast<int64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t ld_u64(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_u64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A 64-bit barrier word that carries a value: the barrier phase mod 2^16 over a 48-bit signed
// value, stored and loaded as one word, so a tag that matches the awaited phase vouches for the
// value beside it.  Equality is enough: a word moves to the next phase only after its reader has
// passed the current one.
__device__ __forceinline__ uint64_t tag_word(uint32_t phase, int64_t v) {
    return ((uint64_t)(phase & 0xffffu) << 48) | ((uint64_t)v & 0xFFFFFFFFFFFFull);
}
__device__ __forceinline__ int64_t word_value(uint64_t w) { return (int64_t)(w << 16) >> 16; }

// Barrier phases count up without bound and wrap at 2^32; a flag is never more than one phase
// ahead of its reader, so "v has not reached phase" is the sign of the 32-bit difference.
__device__ __forceinline__ bool phase_before(uint32_t v, uint32_t phase) {
    return (int32_t)(v - phase) < 0;
}

// Grid barrier with a bounded wait.  Each wave waits for its own stores to complete; then every
// workgroup stores the phase number into its own arrival flag, workgroup 0 polls all flags (one
// thread per flag) and publishes the phase in the generation word, which the others poll.  No
// read-modify-write is involved, so arrivals do not serialise on one address, and only workgroup
// 0 reads the W flags.  A wait longer than timeout_ticks (or another workgroup's abort) sets the
// abort flag; the caller then leaves the kernel, so every wave reaches the end.
__device__ __forceinline__ bool wait_phase(const CoopArgs& C, const uint32_t* p, uint32_t phase) {
    const unsigned long long t0 = (unsigned long long)wall_clock64();
    while (phase_before(ld_u32(p), phase)) {
        if (__hip_atomic_load(C.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            (unsigned long long)wall_clock64() - t0 > C.timeout_ticks) {
            __hip_atomic_store(C.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Wait (bounded, as wait_phase) until *p carries the tag of `phase`; w receives the word.
__device__ __forceinline__ bool wait_tag(const CoopArgs& C, const uint64_t* p, uint32_t phase,
                                         uint64_t& w) {
    const unsigned long long t0 = (unsigned long long)wall_clock64();
    const uint32_t tag = phase & 0xffffu;
    for (;;) {
        w = ld_u64(p);
        if ((uint32_t)(w >> 48) == tag) return true;
        if (__hip_atomic_load(C.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            (unsigned long long)wall_clock64() - t0 > C.timeout_ticks) {
            __hip_atomic_store(C.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// gsum/lsum (optional): an LDS partial sum the waves have added into; thread 0 adds it to the
// global accumulator and waits for that atomic before the workgroup arrives.
__device__ __forceinline__ bool coop_sync(const CoopArgs& C, uint32_t& phase, int* sh_abort,
                                          int64_t* gsum = nullptr, long long* lsum = nullptr) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    ++phase;
    if (lsum && threadIdx.x == 0) {
        const long long v = *lsum;
        *lsum = 0;
        if (v != 0) {
            atomicAdd((unsigned long long*)gsum, (unsigned long long)v);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
    }
    if (blockIdx.x == 0) {
        for (int x = 1 + threadIdx.x; x < (int)gridDim.x; x += blockDim.x)
            if (!wait_phase(C, C.flags + x, phase)) *sh_abort = 1;
        __syncthreads();
        if (threadIdx.x == 0 && *sh_abort == 0) st_u32(C.gen, phase);
    } else if (threadIdx.x == 0) {
        st_u32(C.flags + blockIdx.x, phase);
        if (!wait_phase(C, C.gen, phase)) *sh_abort = 1;
    }
    __syncthreads();
    return *sh_abort == 0;
}

// Round end of R <= 64 replicas by one wavefront, the ladder held in rung order: lane r keeps
// the slot at rung r with that slot's energy, round-trip label and trip count, rung r's inverse
// temperature and rung r's counters.  Exchange partners trade their slot, energy, label and trip
// count through shuffles, so no lane reads another's data through LDS and the wave needs no
// workgroup barrier.  The rules are those of block_best_rung, exchange_pairs and update_labels:
// the same lexicographic best rung, the same pairs, uniforms and acceptance arithmetic, the same
// label transitions; only where each quantity is kept differs.
struct WaveLadder {
    int32_t rep;        // slot at this lane's rung
    int64_t Eb;         // its energy at the last round end
    int32_t lab;        // its label (1: last at the cold end, -1: last at the hot end)
    int32_t trips;      // its completed round trips
    double beta;        // inverse temperature of this rung
    int32_t tried, acc, up, down;   // counters of this rung (tried/acc: pair (r, r+1))
    int64_t bestE;      // best record (uniform over the wave)
    int32_t best_round, best_rung, best_slot, have;
};

__device__ __forceinline__ void wave_ladder_init(WaveLadder& W, const PtInstance& I,
                                                 const int64_t* einit, int lane) {
    const int R = I.R;
    W.rep = lane;
    W.Eb = (einit && lane < R) ? -(einit[lane] / 2) : 0;
    W.lab = (R > 1 && lane == 0) ? 1 : ((R > 1 && lane == R - 1) ? -1 : 0);
    W.trips = 0;
    W.beta = lane < R ? I.beta[lane] : 0.0;
    W.tried = W.acc = W.up = W.down = 0;
    W.bestE = 0;
    W.best_round = W.best_rung = W.best_slot = W.have = 0;
}

// (einit may be null: the energies then come with the first round end.)

// One round end: E is the energy of the slot at this lane's rung after round rnd - 1 (rnd == 0:
// the initial energies).  Updates the best record (returns true if it improved), exchanges the
// pairs of round rnd - 1 and advances the labels; the owner records history and trace.
// uex (optional): the exchange uniforms of round rnd - 1 by lower rung, drawn ahead of time.
__device__ __forceinline__ bool wave_ladder_round(WaveLadder& W, const PtInstance& I, int rnd,
                                                  int64_t E, int lane, bool owner,
                                                  const uint32_t* uex = nullptr) {
    const int R = I.R;
    long long e = lane < R ? (long long)E : LLONG_MAX;
    int r = lane < R ? lane : INT_MAX;
#pragma unroll
    for (int m = PT_WAVE / 2; m > 0; m >>= 1) {   // branch-free lexicographic minimum
        const long long e2 = __shfl_xor(e, m);
        const int r2 = __shfl_xor(r, m);
        const bool take = (e2 < e) | ((e2 == e) & (r2 < r));
        e = take ? e2 : e;
        r = take ? r2 : r;
    }
    const int slot = __shfl(W.rep, r);
    const bool improved = W.have == 0 || e < W.bestE;
    if (owner && rnd > 0) {
        if (lane == 0 && I.hist) I.hist[rnd - 1] = (double)(improved ? e : W.bestE) * I.unit;
        if (I.trace && lane < R) I.trace[(size_t)(rnd - 1) * R + lane] = (double)E * I.unit;
    }
    if (improved) {
        W.bestE = e;
        W.best_round = rnd;
        W.best_rung = r;
        W.best_slot = slot;
        W.have = 1;
    }
    if (rnd > 0 && R > 1) {
        const int k = rnd - 1, par = k & 1;
        const bool lo = lane >= par && ((lane - par) & 1) == 0 && lane + 1 < R;
        const bool hi = lane >= par + 1 && ((lane - 1 - par) & 1) == 0 && lane < R;
        const int pr = lo ? lane + 1 : (hi ? lane - 1 : lane);
        const int rlo = hi ? lane - 1 : lane;
        uint32_t u;
        if (uex) {
            u = uex[rlo < R ? rlo : 0];
        } else {
            const pt_u4 rv = pt_philox((uint32_t)(rlo >> 2), 0u, (uint32_t)k, PT_TAG_EXCH, I.k0, I.k1);
            u = pt_word(rv, rlo & 3);
        }
        const int64_t Ep = __shfl(E, pr);
        const double bp = __shfl(W.beta, pr);
        const int32_t repP = __shfl(W.rep, pr), labP = __shfl(W.lab, pr), tripP = __shfl(W.trips, pr);
        bool ok = false;
        if (lo || hi) {
            const int64_t Elo = lo ? E : Ep, Ehi = lo ? Ep : E;
            const double blo = lo ? W.beta : bp, bhi = lo ? bp : W.beta;
            ok = pt_accept_exchange(blo, bhi, (double)Elo * I.unit, (double)Ehi * I.unit, u);
        }
        if (lo) {
            ++W.tried;
            W.acc += ok ? 1 : 0;
        }
        if (ok) {
            W.rep = repP;
            E = Ep;
            W.lab = labP;
            W.trips = tripP;
        }
        if (lane == 0) {
            if (W.lab == -1) ++W.trips;
            W.lab = 1;
        } else if (lane == R - 1) {
            W.lab = -1;
        }
        W.up += W.lab == 1 ? 1 : 0;
        W.down += W.lab == -1 ? 1 : 0;
    }
    W.Eb = E;
    return improved;
}

// Final bookkeeping of the owner: energies, rungs and labels by slot, the permutation, the best
// record and the counters, in the layouts the per-block round ends leave.
__device__ __forceinline__ void wave_ladder_store(const WaveLadder& W, const PtInstance& I,
                                                  int64_t* Eb, int32_t* rung, int32_t* rep,
                                                  int32_t* label, int64_t* best, int lane) {
    const int R = I.R;
    if (lane < R) {
        Eb[W.rep] = W.Eb;
        rung[W.rep] = lane;
        rep[lane] = W.rep;
        label[W.rep] = W.lab;
        pt_out_arr(I.out, R, PT_OUT_TRIPS)[W.rep] = W.trips;
        pt_out_arr(I.out, R, PT_OUT_UP)[lane] = W.up;
        pt_out_arr(I.out, R, PT_OUT_DOWN)[lane] = W.down;
        if (lane + 1 < R) {
            pt_out_arr(I.out, R, PT_OUT_TRIED)[lane] = W.tried;
            pt_out_arr(I.out, R, PT_OUT_ACC)[lane] = W.acc;
        }
    }
    if (lane == 0) {
        best[0] = W.bestE;
        best[1] = W.best_round;
        best[2] = W.best_rung;
        best[3] = W.best_slot;
        best[4] = W.have;
    }
}

}  // namespace pt_dev
