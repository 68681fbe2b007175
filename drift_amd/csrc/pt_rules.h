// pt_rules.h -- the arithmetic every parallel-tempering decision is made of, shared by the
// gfx950 kernels and the library's host code so both make bit-identical decisions.
//
// Stream contract (key = 64-bit seed, k0 = low word, k1 = high word):
//   init spin i, slot a       philox(ctr = (i>>2, a, 0, TAG_INIT))[i&3] & 1 ? +1 : -1
//   Metropolis, position p    u = philox(ctr = (p>>2, a, t, TAG_METRO))[p&3], t = round*S + sweep
//   exchange, rungs (r, r+1)  u = philox(ctr = (r>>2, 0, round, TAG_EXCH))[r&3]
// "Position" p is the index into the concatenated colour lists, so one Philox call serves four
// same-colour attempts of one slot.  The reference draws from cuRAND / NumPy instead
// (cuda/ising_pt.cu:77,118,144; drift/solvers/parallel_tempering.py:68,93-96).
//
// Acceptance (cuda/ising_pt.cu:118, drift/solvers/parallel_tempering.py:96): flip iff dE <= 0
// or u < 2^32 exp(-beta dE).  Integer couplings compare against a per-rung table
// floor(2^32 exp(-2 beta q)) indexed by q = s_i * local_i (dE = 2q); float couplings evaluate
// expf_det on the device.  Exchange (parallel_tempering.py:104-113): accept iff
// d = (b_r - b_{r+1})(E_r - E_{r+1}) >= 0 or u < 2^32 exp_det(d).
//
// Everything here must be compiled with -ffp-contract=off: each operation rounds once, in
// program order, so host and device agree to the bit.
#pragma once
#include <stdint.h>

#define PT_HD __host__ __device__ __forceinline__

enum : uint32_t { PT_TAG_INIT = 1u, PT_TAG_METRO = 2u, PT_TAG_EXCH = 3u };

struct pt_u4 {
    uint32_t x, y, z, w;
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Multipliers 0xD2511F53 / 0xCD9E8D57,
// Weyl key increments 0x9E3779B9 / 0xBB67AE85.
PT_HD pt_u4 pt_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                      uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t m0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t y0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
        const uint32_t y2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)m1;
        c3 = (uint32_t)m0;
        c0 = y0;
        c2 = y2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return pt_u4{c0, c1, c2, c3};
}

PT_HD uint32_t pt_word(const pt_u4& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

PT_HD double pt_bits_to_double(uint64_t b) {
    union {
        uint64_t u;
        double d;
    } c;
    c.u = b;
    return c.d;
}

PT_HD float pt_bits_to_float(uint32_t b) {
    union {
        uint32_t u;
        float f;
    } c;
    c.u = b;
    return c.f;
}

// exp in double: Cody-Waite split of ln 2 (high part with 32 trailing zero bits, so n*ln2_hi is
// exact), degree-13 Taylor polynomial in Horner form (coefficients 1/k! correctly rounded), then
// an exact power-of-two scale.  Arguments below -700 give 0 (only ratios >= 2^-32 matter here).
PT_HD double pt_exp_det(double x) {
    if (!(x > -700.0)) return 0.0;
    if (x > 700.0) x = 700.0;
    const double nf = floor(x * 0x1.71547652b82fep+0 + 0.5);
    const double r = (x - nf * 0x1.62e42fee00000p-1) - nf * 0x1.a39ef35793c76p-33;
    double p = 0x1.6124613a86d09p-33;
    p = p * r + 0x1.1eed8eff8d898p-29;
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
    const int n = (int)nf;
    return p * pt_bits_to_double((uint64_t)(int64_t)(n + 1023) << 52);
}

// exp in float: Cody-Waite with ln2 = 0x1.63p-1 - 0x1.bd0106p-13 (the high part has 8 significant
// bits, so n*hi is exact for |n| <= 127), degree-7 Taylor polynomial.  Arguments below -87 give 0,
// so no result is subnormal.
PT_HD float pt_expf_det(float x) {
    if (!(x > -87.0f)) return 0.0f;
    if (x > 87.0f) x = 87.0f;
    const float nf = floorf(x * 0x1.715476p+0f + 0.5f);
    const float r = (x - nf * 0x1.63p-1f) - nf * -0x1.bd0106p-13f;
    float p = 0x1.a01a02p-13f;
    p = p * r + 0x1.6c16c2p-10f;
    p = p * r + 0x1.111112p-7f;
    p = p * r + 0x1.555556p-5f;
    p = p * r + 0x1.555556p-3f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int n = (int)nf;
    return p * pt_bits_to_float((uint32_t)(n + 127) << 23);
}

// Threshold for an uphill integer move dE = 2q (q >= 1) at inverse temperature beta*scale.
PT_HD uint32_t pt_threshold_int(double beta_scaled, int32_t q) {
    const double x = -2.0 * beta_scaled * (double)q;
    const double v = pt_exp_det(x) * 4294967296.0;
    return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
}

// Float-coupling Metropolis decision for q = s_i * local_i (dE = 2q).
PT_HD bool pt_accept_float(float beta, float q, uint32_t u) {
    if (q <= 0.0f) return true;
    const float dE = 2.0f * q;
    const float x = -(beta * dE);
    const float p = pt_expf_det(x);
    const uint32_t thr = p >= 1.0f ? 0xFFFFFFFFu : (uint32_t)(p * 4294967296.0f);
    return u < thr;
}

// Replica-exchange decision for the pair (rung r at b_lo, rung r+1 at b_hi).
PT_HD bool pt_accept_exchange(double b_lo, double b_hi, double e_lo, double e_hi, uint32_t u) {
    const double d = (b_lo - b_hi) * (e_lo - e_hi);
    if (d >= 0.0) return true;
    return (double)u < pt_exp_det(d) * 4294967296.0;
}

// Float energies are carried as int64 multiples of 2^-36 so that sums are associative (the
// same in any reduction order) and therefore identical on every engine and on the CPU.
// v in 2^-36 units, rounded to the nearest integer (ties to even): adding 1.5 * 2^52 to the exact
// v * 2^36 leaves that integer in the low mantissa bits, read back by subtracting the constant's
// bits.  Exact for |v * 2^36| < 2^51, which the host's bound on float local fields (|field| <
// 2^14, so |2 q| < 2^15) ensures.
PT_HD int64_t pt_fix36d(double v) {
    union {
        double d;
        int64_t i;
    } c;
    c.d = v * 68719476736.0 + 6755399441055744.0;
    return c.i - (int64_t)0x4338000000000000ll;
}

// Float couplings: the local field h_i + sum_j J_ij s_j of float32 values is summed in double, in
// row order.  A double holds the sum of float32 values whose magnitudes lie within 2^29 of one
// another exactly, so the field is (for such couplings) the exact real sum, and the energy change
// of a flip, pt_fix36d(2 q) with q = s_i * field, is exact in 2^-36 units for any coupling that is
// a multiple of 2^-36 (every float32 of magnitude at least 2^-13).  The tracked energies are
// therefore the energies of the spins, not a sum of rounded float fields that drifts with the
// number of accepted flips.  The Metropolis decision takes the field rounded once to float.

#define PT_FIX36_UNIT 0x1p-36
