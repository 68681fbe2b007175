// pt_engine.hip -- gfx950 kernels of the parallel-tempering engine.
//
// Two engines share one Markov chain (so their results are bit-identical):
//
//  * LDS engine (pt_lds_kernel): one workgroup owns a whole instance -- every replica's int8 spins
//    live in LDS as [position][slot], so neighbour gathers for the slots of one position are
//    contiguous bytes.  The workgroup runs all rounds itself: colour steps are separated by
//    s_barrier, and the exchange, best tracking and round-trip bookkeeping happen in LDS with no
//    host round trip.  A global work counter hands instances (largest first) to persistent
//    workgroups, which is how thousands of uneven instances run in one launch.
//
//  * Global engine (g_* kernels): for instances whose replicas do not fit one workgroup's LDS.
//    Spins live in HBM as [position][slot] (a wave reads the 64 slots of a neighbour as one
//    64-byte row); each launch updates one colour class of all replicas across the whole chip.
//    A small round-end launch reduces the energy changes, tracks the best state, copies it, and
//    makes the exchange decisions on the device.
//
// Reference counterpart: cuda/ising_pt.cu:70-146 (init_kernel, sweep_kernel, swap_kernel), which
// used one warp per replica, float spins in global memory, fp64 local fields and cuRAND.
#include "pt_engine.h"
#include "pt_rules.h"
#include "pt_device.h"

#include <algorithm>
#include <climits>


namespace {

using namespace pt_dev;

// ------------------------------------------------------------------------------------------
// LDS engine
// ------------------------------------------------------------------------------------------
struct LShared {
    long long bestE;
    int have, bround, brung, bslot, improved, inst, abort;
};

// Split engine mailbox word: a round tag (16 bits) over 32 bits of an energy.  A slot's energy
// travels as two such words (low and high half), each checked against the awaited tag, so a
// reader never combines halves of different rounds.  Buffers alternate by round parity: a
// workgroup writes round r + 2 only after reading every slot's round r + 1 words, each written
// after its writer had read round r, so equal tags two rounds apart never meet.
__device__ __forceinline__ uint64_t pt_split_word(uint32_t tag, uint32_t half) {
    return ((uint64_t)(tag & 0xffffu) << 48) | (uint64_t)half;
}

// Split engine, after round rnd's sweeps: publish this group's slots' energies and gather every
// slot's into E (wave 0 reads; one lane per slot).  False when a wait exceeded its bound or
// another workgroup gave up (the abort flag is then set, and every workgroup leaves).
// Instances of many groups and slots (ng * R >= PT_SPLIT_HOP2) gather in two hops: group 0 alone
// polls every slot's words, then stores the round's tag into the round word (box[4 Rp + parity]);
// the other groups wait on that one word and then read each slot's words once.  Every group
// polling every slot would put ng * 2R coherent loads per round on the same few cache lines
// (131 000 for R = 256 in 256 groups).  The tags are checked in both cases, so a group never takes
// another round's energy whatever order the stores become visible in.
#define PT_SPLIT_HOP2 1024
__device__ bool split_gather(const PtInstance& I, const SplitArgs& A, int rnd, int grp, int a0,
                             int rg, int64_t* E, LShared* sc) {
    const int tid = threadIdx.x, R = I.R;
    uint64_t* box = I.box + (size_t)(rnd & 1) * 2 * I.Rp;
    uint64_t* round_word = I.box + (size_t)4 * I.Rp + (rnd & 1);
    const uint32_t tag = (uint32_t)(rnd + 1) & 0xffffu;
    const bool hop2 = I.ng * R >= PT_SPLIT_HOP2;
    if (tid < rg && a0 + tid < R) {
        const uint64_t e = (uint64_t)E[a0 + tid];
        st_u64(box + 2 * (a0 + tid), pt_split_word(tag, (uint32_t)e));
        st_u64(box + 2 * (a0 + tid) + 1, pt_split_word(tag, (uint32_t)(e >> 32)));
    }
    if (tid < PT_WAVE) {
        bool ok = true;
        const unsigned long long t0 = (unsigned long long)wall_clock64();
        auto give_up = [&]() {
            if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                (unsigned long long)wall_clock64() - t0 > A.timeout_ticks) {
                __hip_atomic_store(A.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return true;
            }
            return false;
        };
        if (hop2 && grp != 0) {
            // the round word first: group 0 stores it once it has read every slot's words
            while ((uint32_t)(ld_u64(round_word) >> 48) != tag) {
                if (give_up()) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        for (int x = tid; x < R && ok; x += PT_WAVE) {
            if (x >= a0 && x < a0 + rg) continue;
            for (;;) {
                const uint64_t lo = ld_u64(box + 2 * x), hi = ld_u64(box + 2 * x + 1);
                if ((uint32_t)(lo >> 48) == tag && (uint32_t)(hi >> 48) == tag) {
                    E[x] = (int64_t)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull));
                    break;
                }
                if (give_up()) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (hop2 && grp == 0 && ok && tid == 0) st_u64(round_word, pt_split_word(tag, 0u));
        if (!ok) sc->abort = 1;
    }
    __syncthreads();
    return sc->abort == 0;
}

// Split engine with parts: colour class [cs, ce) covers the 32-position words w0 .. w0 + nwc - 1
// (word w: positions 32 w .. 32 w + 31); part q of np owns the run [ow0, ow1) of them, the runs
// as equal as integer division makes them, so every part owns a word of any class of at least
// np words.  Each word of the class has exactly one owner, and holds at least one position of
// the class.
struct PartRun {
    int w0, nwc, ow0, ow1;
};
__device__ __forceinline__ PartRun part_run(int cs, int ce, int np, int q) {
    PartRun r;
    r.w0 = cs >> 5;
    r.nwc = ((ce - 1) >> 5) + 1 - r.w0;
    r.ow0 = r.w0 + (int)(((int64_t)q * r.nwc) / np);
    r.ow1 = r.w0 + (int)(((int64_t)(q + 1) * r.nwc) / np);
    return r;
}

// Sign bits of four int8 spins held in a word (byte b: position 4 i + b) as bits 0..3, and back
// (bit set: spin -1, else +1).
__device__ __forceinline__ uint32_t sign_nibble(uint32_t x) {
    return ((x >> 7) & 1u) | ((x >> 14) & 2u) | ((x >> 21) & 4u) | ((x >> 28) & 8u);
}
__device__ __forceinline__ uint32_t nibble_spins(uint32_t b) {
    const uint32_t e = (b & 1u) | ((b & 2u) << 7) | ((b & 4u) << 14) | ((b & 8u) << 21);
    return 0x01010101u | (e * 0xFEu);
}

// Split engine with parts: a colour class of at most two positions per thread of the workgroup
// is swept whole by every part of the group -- the same positions, streams and neighbours, so the
// same result in every part -- instead of being traded: one pass over it costs less than a trade.
// Every part keeps the class's energy change apart (drep), and the group's total counts it once.
__device__ __forceinline__ bool part_replicated(int cs, int ce) { return ce - cs <= 2 * (int)blockDim.x; }

// Split engine with parts, after colour step (sweep t, colour class [cs, ce), traded): publish
// this part's words of the class for every live slot of the group under the step tag stag, raise
// this part's flag to stag, wait until the group's other parts have raised theirs, then read the
// class's other words into the LDS spins (only the positions of the class are written).  The
// flags only spare the readers from polling every word: each word carries its tag, and a word
// read before its store is visible is read again until it carries stag or a later step.  A later
// step is accepted because a word holding the end of one class and the start of the next is
// written again by the next step's owner, with the first class's bits unchanged.  The words
// alternate between two buffers by sweep parity: a part writes sweep t + 2's words only after
// reading every part's words of the largest class of sweep t + 1 (never replicated unless every
// class is, and every part owns a word of it: the host keeps np at most its word count), each
// written after its writer had read all of sweep t, so no word is overwritten while a reader
// still needs it.  ride (the round's last traded step): each part also publishes its slots'
// energy changes on traded classes since the start (dpart, two round-tagged words each in the
// first half of pbox, by round parity), and after the flags every part sums the group's into
// gpart, so the round end needs no hop inside the group.  False when a wait exceeded its bound
// or another workgroup gave up (the abort flag is then set, and every workgroup leaves).
__device__ bool part_trade(const PtInstance& I, const SplitArgs& A, int cs, int ce, int np,
                           int part, int grp, int a0, int rg, int8_t* sp, uint32_t t,
                           uint32_t stag, LShared* sc, bool ride, int rnd, const int64_t* dpart,
                           int64_t* gpart) {
    const int tid = threadIdx.x, nt = blockDim.x, n = I.n, NW = I.nw;
    const int rl = I.R - a0 < rg ? I.R - a0 : rg;   // live slots of the group
    const PartRun pr = part_run(cs, ce, np, part);
    const int own = pr.ow1 - pr.ow0;
    uint64_t* mir = I.pbits + ((size_t)(t & 1u) * I.Rp + a0) * NW;   // [slot][word]
    uint64_t* flags = I.pflag + (size_t)grp * np * PT_PFLAG_STRIDE;
    for (int idx = tid; idx < own * rl; idx += nt) {
        const int s = idx % rl, w = pr.ow0 + idx / rl, pb = w << 5;
        uint32_t bits = 0;
        if (rg == 1 && pb + 32 <= n) {
            const uint4 x0 = ((const uint4*)(sp + pb))[0], x1 = ((const uint4*)(sp + pb))[1];
            bits = sign_nibble(x0.x) | (sign_nibble(x0.y) << 4) | (sign_nibble(x0.z) << 8) |
                   (sign_nibble(x0.w) << 12) | (sign_nibble(x1.x) << 16) |
                   (sign_nibble(x1.y) << 20) | (sign_nibble(x1.z) << 24) |
                   (sign_nibble(x1.w) << 28);
        } else {
            for (int b = 0; b < 32 && pb + b < n; ++b)
                bits |= (uint32_t)(sp[(size_t)(pb + b) * rg + s] < 0) << b;
        }
        st_u64(mir + (size_t)s * NW + w, ((uint64_t)stag << 32) | bits);
    }
    const uint32_t rtag = (uint32_t)(rnd + 1) & 0xffffu;
    uint64_t* pb = I.pbox + (size_t)(rnd & 1) * 2 * np * I.Rp;   // [part][slot][2]
    if (ride && tid < rl) {
        const uint64_t e = (uint64_t)dpart[tid];
        uint64_t* d = pb + 2 * ((size_t)part * I.Rp + a0 + tid);
        st_u64(d, pt_split_word(rtag, (uint32_t)e));
        st_u64(d + 1, pt_split_word(rtag, (uint32_t)(e >> 32)));
        gpart[tid] = 0;
    }
    __syncthreads();
    if (tid == 0) st_u64(flags + (size_t)part * PT_PFLAG_STRIDE, (uint64_t)stag);
    auto give_up = [&](unsigned long long t0) {
        if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            (unsigned long long)wall_clock64() - t0 > A.timeout_ticks) {
            __hip_atomic_store(A.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
        return false;
    };
    if (tid < np && tid != part) {
        const uint64_t* f = flags + (size_t)tid * PT_PFLAG_STRIDE;
        const unsigned long long t0 = (unsigned long long)wall_clock64();
        while ((int32_t)((uint32_t)ld_u64(f) - stag) < 0) {
            if (give_up(t0)) {
                sc->abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    if (sc->abort) return false;
    bool ok = true;
    if (ride)
        for (int idx = tid; idx < np * rl && ok; idx += nt) {
            const int q = idx / rl, x = idx - q * rl;
            int64_t v = dpart[x];
            if (q != part) {
                const uint64_t* w2 = pb + 2 * ((size_t)q * I.Rp + a0 + x);
                const unsigned long long t0 = (unsigned long long)wall_clock64();
                for (;;) {
                    const uint64_t lo = ld_u64(w2), hi = ld_u64(w2 + 1);
                    if ((uint32_t)(lo >> 48) == rtag && (uint32_t)(hi >> 48) == rtag) {
                        v = (int64_t)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull));
                        break;
                    }
                    if (give_up(t0)) {
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (ok && v != 0) atomicAdd((unsigned long long*)&gpart[x], (unsigned long long)v);
        }
    const int other = pr.nwc - own;
    for (int idx = tid; idx < other * rl; idx += nt) {
        const int s = idx % rl;
        int w = pr.w0 + idx / rl;
        if (w >= pr.ow0) w += own;
        const uint64_t* src = mir + (size_t)s * NW + w;
        const unsigned long long t0 = (unsigned long long)wall_clock64();
        uint64_t v = ld_u64(src);
        while ((int32_t)((uint32_t)(v >> 32) - stag) < 0) {
            if (give_up(t0)) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            v = ld_u64(src);
        }
        if (!ok) break;
        const uint32_t bits = (uint32_t)v;
        const int pb = w << 5;
        const int lo = cs > pb ? cs : pb, hi = ce < pb + 32 ? ce : pb + 32;
        if (rg == 1 && lo == pb && hi == pb + 32) {
            uint4 x0, x1;
            x0.x = nibble_spins(bits & 15u);         x0.y = nibble_spins((bits >> 4) & 15u);
            x0.z = nibble_spins((bits >> 8) & 15u);  x0.w = nibble_spins((bits >> 12) & 15u);
            x1.x = nibble_spins((bits >> 16) & 15u); x1.y = nibble_spins((bits >> 20) & 15u);
            x1.z = nibble_spins((bits >> 24) & 15u); x1.w = nibble_spins(bits >> 28);
            ((uint4*)(sp + pb))[0] = x0;
            ((uint4*)(sp + pb))[1] = x1;
        } else {
            for (int p = lo; p < hi; ++p)
                sp[(size_t)p * rg + s] = ((bits >> (p & 31)) & 1u) ? (int8_t)-1 : (int8_t)1;
        }
    }
    if (!ok) sc->abort = 1;
    __syncthreads();
    return sc->abort == 0;
}

// Split engine with parts, after round rnd's sweeps: a slot's energy is its initial energy plus
// its group's changes -- on traded classes the sum over the group's parts (gpart, gathered at
// the round's last traded step; a group in one part, or trading nothing, has only its own,
// dpart) and on the classes every part sweeps whole (drep, the same in every part).  Every part
// of a group so holds its slots' energies; each publishes them (two round-tagged words per slot,
// in the second half of pbox by round parity, under its part index), and every workgroup reads
// the other groups' slots from the part of its own index.  A part writes round r + 2's words
// only after reading every slot's words of round r + 1 from the parts of its index, each written
// after its writer had read round r, so two buffers suffice.  False when a wait exceeded its
// bound.
__device__ bool part_gather(const PtInstance& I, const SplitArgs& A, int rnd, int part, int a0,
                            int rg, int64_t* E, const int64_t* e0l, const int64_t* dpart,
                            const int64_t* gpart, const int64_t* drep, bool traded,
                            LShared* sc) {
    const int tid = threadIdx.x, nt = blockDim.x, R = I.R, Rp = I.Rp, np = I.np;
    const uint32_t tag = (uint32_t)(rnd + 1) & 0xffffu;
    uint64_t* gb = I.pbox + (size_t)4 * np * Rp + ((size_t)(rnd & 1) * np + part) * 2 * Rp;
    const int rl = R - a0 < rg ? R - a0 : rg;
    if (tid < rl) {
        const int64_t e = e0l[a0 + tid] + (np > 1 && traded ? gpart[tid] : dpart[tid]) + drep[tid];
        E[a0 + tid] = e;
        st_u64(gb + 2 * (a0 + tid), pt_split_word(tag, (uint32_t)(uint64_t)e));
        st_u64(gb + 2 * (a0 + tid) + 1, pt_split_word(tag, (uint32_t)((uint64_t)e >> 32)));
    }
    bool ok = true;
    for (int x = tid; x < R && ok; x += nt) {
        if (x >= a0 && x < a0 + rl) continue;
        const uint64_t* w2 = gb + 2 * x;
        const unsigned long long t0 = (unsigned long long)wall_clock64();
        for (;;) {
            const uint64_t lo = ld_u64(w2), hi = ld_u64(w2 + 1);
            if ((uint32_t)(lo >> 48) == tag && (uint32_t)(hi >> 48) == tag) {
                E[x] = (int64_t)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull));
                break;
            }
            if (__hip_atomic_load(A.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                (unsigned long long)wall_clock64() - t0 > A.timeout_ticks) {
                __hip_atomic_store(A.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (!ok) sc->abort = 1;
    __syncthreads();
    return sc->abort == 0;
}

// One instance (SPLIT = false: all its slots; SPLIT: the rg slots of group grp, the other groups
// of the instance running at the same time in other workgroups).
// INIT: draw the initial spins and energies and leave them in I.img / I.e0 (the launch before the
// timed one).  Otherwise: load them and run every round.
// Split groups run the same chain: the sweeps of a slot depend only on its own spins and on the
// rung it holds; after a round's sweeps every group reads every slot's energy from the mailboxes,
// and every group then takes the same round-end decisions (best record, exchanges, labels) from
// the same inputs.  The group holding the best slot copies its state; group 0 writes the history,
// the trace and the statistics.
// CL (split engine, int8 couplings): the instance's row pointers, 16-bit records and field are
// copied into LDS after the ladder arrays, and the sweeps gather from there (quad_update_cl).
// PARTS (split engine): work item wi is group wi / np, part wi % np; the workgroup holds the
// group's whole state and sweeps its part's words of each colour class (part_run), trading the
// class's new spins with the group's other parts after every colour step (part_trade, which at
// the round's last traded step also sums the group's energy changes) and the slots' energies
// after every round (part_gather).  Every part then takes the same
// round-end decisions; the parts of the group holding a new best slot each copy their share of
// its positions.
template <int WM, bool INIT, bool SPLIT, bool CL = false, bool PARTS = false>
__device__ void lds_instance(const PtInstance& I, int wi, char* smem, LShared* sc,
                             long long* scrE, int* scrR, const SplitArgs& SA) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = I.n, R = I.R, Rp = I.Rp, k = I.k;
    const int np = PARTS ? I.np : 1;
    const int part = PARTS ? wi % np : 0;
    const int grp = PARTS ? wi / np : wi;
    const int rg = SPLIT ? I.rg : Rp;          // slots this workgroup holds
    const int a0 = SPLIT ? grp * rg : 0;       // the first of them
    const bool owner = !SPLIT || (grp == 0 && part == 0);   // writes history, trace, statistics
    // positions this part copies when its group holds a new best slot, and writes at the end
    const int pp0 = (int)(((int64_t)part * n) / np), pp1 = (int)(((int64_t)(part + 1) * n) / np);
    size_t off = ((size_t)n * rg + 15) & ~(size_t)15;
    int8_t* sp = (int8_t*)smem;
    int64_t* E = (int64_t*)(smem + off);      off += 8 * (size_t)Rp;
    int64_t* tried = (int64_t*)(smem + off);  off += 8 * (size_t)Rp;
    int64_t* accd = (int64_t*)(smem + off);   off += 8 * (size_t)Rp;
    int64_t* up = (int64_t*)(smem + off);     off += 8 * (size_t)Rp;
    int64_t* down = (int64_t*)(smem + off);   off += 8 * (size_t)Rp;
    int32_t* rung = (int32_t*)(smem + off);   off += 4 * (size_t)Rp;
    int32_t* rep = (int32_t*)(smem + off);    off += 4 * (size_t)Rp;
    int32_t* label = (int32_t*)(smem + off);  off += 4 * (size_t)Rp;
    int32_t* trips = (int32_t*)(smem + off);  off += 4 * (size_t)Rp;
    float* beta_f = (float*)(smem + off);     off += 4 * (size_t)Rp;
    off = (off + 15) & ~(size_t)15;
    int32_t* cptr = (int32_t*)(smem + off);   off += ((size_t)4 * (k + 1) + 15) & ~(size_t)15;
    uint32_t* thr_l = (uint32_t*)(smem + off);
    const bool thr_in_lds = SPLIT && WM != PT_WM_F32 && I.thr_lds != 0;
    if (thr_in_lds) off += ((size_t)4 * rg * I.TQ + 15) & ~(size_t)15;
    int64_t* e0l = (int64_t*)(smem + off);    // PARTS: initial energy of every slot
    int64_t* dpart = e0l + Rp;                // PARTS: this part's change of each of its slots on
                                              // traded classes, since the start
    int64_t* gpart = dpart + rg;              // PARTS: the group's sum of dpart
    int64_t* drep = gpart + rg;               // PARTS: the change on classes every part sweeps
    if constexpr (PARTS) off += ((size_t)8 * (Rp + 3 * rg) + 15) & ~(size_t)15;
    int32_t* rpl = (int32_t*)(smem + off);    off += ((size_t)4 * (n + 1) + 15) & ~(size_t)15;
    uint16_t* recl = (uint16_t*)(smem + off); off += ((size_t)2 * I.nnz + 15) & ~(size_t)15;
    int32_t* hl = I.h ? (int32_t*)(smem + off) : nullptr;

    const int al = tid % rg, qo = tid / rg, qs = nt / rg;
    const int a = a0 + al;                     // the slot this thread sweeps
    const int img16 = (int)(pt_lds_img_bytes(n, rg) >> 4);   // the spins as 16-byte words
    int8_t* img = SPLIT ? I.img + (size_t)grp * pt_lds_img_bytes(n, rg) : I.img;
    if constexpr (INIT) {
        // --- initial spins from the stream and their energies, out to the image ---------------
        for (int idx = tid; idx < n * rg; idx += nt) {
            const int p = idx / rg, s = idx - p * rg;
            sp[idx] = a0 + s < R ? init_spin(I, p, a0 + s) : (int8_t)0;
        }
        for (int x = tid; x < rg; x += nt) E[a0 + x] = 0;
        __syncthreads();
        int64_t acc = 0;
        if (a < R)
            for (int p = qo; p < n; p += qs)
                acc += site_energy<WM>(p, al, rg, sp, I.row_ptr, I.edge, I.h);
        if (acc != 0) atomicAdd((unsigned long long*)&E[a], (unsigned long long)acc);
        __syncthreads();
        for (int x = tid; x < img16; x += nt) ((uint4*)img)[x] = ((const uint4*)sp)[x];
        for (int x = tid; x < rg; x += nt)
            I.e0[a0 + x] = WM != PT_WM_F32 ? -(E[a0 + x] / 2) : E[a0 + x];
        return;
    }
    // --- the initial state from the image, bookkeeping ---------------------------------------
    for (int x = tid; x < img16; x += nt) ((uint4*)sp)[x] = ((const uint4*)img)[x];
    for (int x = tid; x < Rp; x += nt) {
        E[x] = x < R ? I.e0[x] : 0; tried[x] = 0; accd[x] = 0; up[x] = 0; down[x] = 0; trips[x] = 0;
        if constexpr (PARTS) e0l[x] = E[x];
        rung[x] = x;
        rep[x] = x;
        label[x] = (R > 1 && x == 0) ? 1 : ((R > 1 && x == R - 1) ? -1 : 0);
        beta_f[x] = x < R ? (float)I.beta[x] : 0.0f;
    }
    if (tid == 0) {
        sc->have = 0; sc->bestE = 0; sc->bround = 0; sc->brung = 0; sc->bslot = 0; sc->abort = 0;
    }
    for (int x = tid; x <= k; x += nt) cptr[x] = I.color_ptr[x];
    if constexpr (PARTS)
        for (int x = tid; x < rg; x += nt) dpart[x] = gpart[x] = drep[x] = 0;
    if constexpr (CL) {
        for (int x = tid; x <= n; x += nt) rpl[x] = I.row_ptr[x];
        for (int x = tid; x < I.nnz; x += nt) recl[x] = I.crec[x];
        if (hl)
            for (int x = tid; x < n; x += nt) hl[x] = ((const int32_t*)I.h)[x];
    }
    __syncthreads();

    // PARTS: the last colour class the parts trade (-1: none), whose step in a round's last sweep
    // carries the parts' energy changes
    int last_tc = -1;
    if constexpr (PARTS)
        if (np > 1)
            for (int c = 0; c < k; ++c)
                if (cptr[c + 1] > cptr[c] && !part_replicated(cptr[c], cptr[c + 1])) last_tc = c;

    // R <= 64: the round end is wave 0's (WaveLadder), with one workgroup barrier per round.  The
    // ladder stays in the LDS arrays between round ends, so no registers are held over the sweeps.
    const bool fast = R <= PT_WAVE;
    // the best record sits in LDS after every round end's barrier, so every thread takes the same
    // decision to stop
    int rounds_run = I.n_rounds;
    auto reached = [&]() { return I.has_target && sc->have && sc->bestE <= I.target_units; };
    // copy of the best slot's state: the whole instance, or (split) only by the group holding it,
    // with coherent stores so that a later group's copy is the one memory keeps
    auto copy_best = [&](int slot) {
        if constexpr (SPLIT) {
            if (slot >= a0 && slot < a0 + rg)
                for (int p = pp0 + tid; p < pp1; p += nt)
                    st_spin<true>(I.best_s + p, sp[(size_t)p * rg + slot - a0]);
        } else {
            for (int p = tid; p < n; p += nt) I.best_s[p] = sp[(size_t)p * Rp + slot];
        }
    };

    for (int rnd = -1; rnd < I.n_rounds; ++rnd) {
        if (rnd >= 0) {
            // --- sweeps of this round --------------------------------------------------------
            // Every thread runs the same loop nest so all waves meet the same barriers; padded
            // slots (a >= R) only skip the work.
            int64_t dsum = 0;
            const bool live = a < R;
            const int g = live ? rung[a] : 0;
            const float bf = beta_f[g];
            const uint32_t* thr_row = WM == PT_WM_F32 ? nullptr : I.thr + (size_t)g * I.TQ;
            if (thr_in_lds) {
                // the threshold rows of this round's rungs of the group's slots, one copy per
                // round instead of a dependent load from HBM in every colour step
                const int TQ = I.TQ;
                for (int idx = tid; idx < rg * TQ; idx += nt) {
                    const int s = idx / TQ, q = idx - s * TQ;
                    thr_l[idx] = a0 + s < R ? I.thr[(size_t)rung[a0 + s] * TQ + q] : 0u;
                }
                __syncthreads();
                thr_row = thr_l + (size_t)al * TQ;
            }
            for (int j = 0; j < I.S; ++j) {
                const uint32_t t = (uint32_t)(rnd * I.S + j);
                int qbase = 0;   // the class's first entry in I.qorder
                for (int c = 0; c < k; ++c) {
                    const int cs = cptr[c], ce = cptr[c + 1];
                    const int nqc = ce > cs ? ((ce - 1) >> 2) - (cs >> 2) + 1 : 0;
                    if constexpr (CL) {
                        if (live && ce > cs) {
                            const int q1 = (ce - 1) >> 2;
                            for (int qq = (cs >> 2) + qo; qq <= q1; qq += qs) {
                                const pt_u4 rv = pt_philox((uint32_t)qq, (uint32_t)a, t,
                                                           PT_TAG_METRO, I.k0, I.k1);
                                dsum += quad_update_cl(qq, cs, ce, al, rg, sp, rpl, recl, hl, rv,
                                                       thr_row, I.TQ);
                            }
                        }
                    } else if (live && ce > cs) {
                        int qb = cs >> 2, q1 = (ce - 1) >> 2;
                        bool whole = false;   // every part sweeps the class
                        if constexpr (PARTS) {
                            if (np > 1 && part_replicated(cs, ce)) {
                                whole = true;
                            } else {
                                // this part's quads: those of its words (eight quads each)
                                const PartRun pr = part_run(cs, ce, np, part);
                                qb = qb > pr.ow0 * 8 ? qb : pr.ow0 * 8;
                                q1 = q1 < pr.ow1 * 8 - 1 ? q1 : pr.ow1 * 8 - 1;
                            }
                        }
                        // int8 couplings: the class's quads longest row first where the instance
                        // has an order (not with parts, whose runs are ranges of words)
                        const int32_t* qord = nullptr;
                        if constexpr (WM == PT_WM_I8 && !PARTS)
                            if (I.qorder) qord = I.qorder + qbase;
                        auto quads = [&](auto gw) {
                            auto one = [&](int qq) {
                                const pt_u4 rv = pt_philox((uint32_t)qq, (uint32_t)a, t,
                                                           PT_TAG_METRO, I.k0, I.k1);
                                const int64_t d = quad_update<WM, decltype(gw)::value>(
                                    qq, cs, ce, n, al, rg, sp, I.row_ptr, I.edge, I.ell, I.h,
                                    I.ell_rest != 0, rv, thr_row, I.TQ, bf);
                                // classes every part sweeps: straight into drep (they are small)
                                if (PARTS && whole) {
                                    if (d != 0)
                                        atomicAdd((unsigned long long*)&drep[al], (unsigned long long)d);
                                } else {
                                    dsum += d;
                                }
                            };
                            if constexpr (WM == PT_WM_I8 && !PARTS) {
                                if (qord) {
                                    for (int i = qo; i <= q1 - qb; i += qs) one(qord[i]);
                                    return;
                                }
                            }
                            for (int qq = qb + qo; qq <= q1; qq += qs) one(qq);
                        };
                        switch (I.gq) {
                            case 2: quads(std::integral_constant<int, 2>{}); break;
                            case 3: quads(std::integral_constant<int, 3>{}); break;
                            case 6:
                                if constexpr (WM == PT_WM_I8) quads(std::integral_constant<int, 6>{});
                                break;
                            default: quads(std::integral_constant<int, 4>{}); break;
                        }
                    }
                    qbase += nqc;
                    const bool ride = PARTS && c == last_tc && j == I.S - 1;
                    if (ride) {
                        // the traded classes' changes so far go out with this step's words
                        if (dsum != 0)
                            atomicAdd((unsigned long long*)&dpart[al], (unsigned long long)dsum);
                        dsum = 0;
                    }
                    __syncthreads();
                    if constexpr (PARTS) {
                        if (np > 1 && ce > cs && !part_replicated(cs, ce) &&
                            !part_trade(I, SA, cs, ce, np, part, grp, a0, rg, sp, t,
                                        t * (uint32_t)k + (uint32_t)c + 1u, sc, ride, rnd, dpart,
                                        gpart))
                            return;
                    }
                }
            }
            if (dsum != 0)
                atomicAdd((unsigned long long*)(PARTS ? &dpart[al] : &E[a]), (unsigned long long)dsum);
            // split: the best-state copy of the last round end is complete before the energies
            // that let other groups pass this round go out
            if constexpr (SPLIT) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            if constexpr (PARTS) {
                if (!part_gather(I, SA, rnd, part, a0, rg, E, e0l, dpart, gpart, drep,
                                 last_tc >= 0, sc))
                    return;
            } else if constexpr (SPLIT) {
                if (!split_gather(I, SA, rnd, grp, a0, rg, E, sc)) return;
            }
        }

        if (fast) {
            if (tid < PT_WAVE) {
                const int r = tid;
                WaveLadder W;
                W.rep = r < R ? rep[r] : r;
                W.Eb = 0;
                W.lab = r < R ? label[W.rep] : 0;
                W.trips = r < R ? trips[W.rep] : 0;
                W.beta = r < R ? I.beta[r] : 0.0;
                W.tried = r < R ? (int32_t)tried[r] : 0;
                W.acc = r < R ? (int32_t)accd[r] : 0;
                W.up = r < R ? (int32_t)up[r] : 0;
                W.down = r < R ? (int32_t)down[r] : 0;
                W.bestE = sc->bestE;
                W.have = sc->have;
                W.best_round = sc->bround;
                W.best_rung = sc->brung;
                W.best_slot = sc->bslot;
                const int64_t Ecur = r < R ? E[W.rep] : 0;
                const bool imp = wave_ladder_round(W, I, rnd + 1, Ecur, r, owner);
                if (r < R) {
                    rep[r] = W.rep;
                    rung[W.rep] = r;
                    label[W.rep] = W.lab;
                    trips[W.rep] = W.trips;
                    tried[r] = W.tried;
                    accd[r] = W.acc;
                    up[r] = W.up;
                    down[r] = W.down;
                }
                if (r == 0) {
                    sc->improved = imp;
                    sc->bestE = W.bestE;
                    sc->have = W.have;
                    sc->bround = W.best_round;
                    sc->brung = W.best_rung;
                    sc->bslot = W.best_slot;
                }
            }
            __syncthreads();
            if (sc->improved) {
                copy_best(sc->bslot);
                __syncthreads();   // read before the next round's sweeps change the spins
            }
            if (reached()) {
                rounds_run = rnd + 1;
                break;
            }
            continue;
        }
        // --- best tracking at round granularity -------------------------------------------
        int64_t mE;
        int32_t mR;
        block_best_rung(E, rep, R, scrE, scrR, mE, mR);
        const bool improved = !sc->have || mE < sc->bestE;
        const int slot = rep[mR];
        __syncthreads();
        if (tid == 0) {
            if (improved) {
                sc->bestE = mE; sc->bround = rnd + 1; sc->brung = mR; sc->bslot = slot; sc->have = 1;
            }
            if (rnd >= 0 && I.hist && owner) I.hist[rnd] = (double)sc->bestE * I.unit;
        }
        if (improved) copy_best(slot);
        if (rnd >= 0 && I.trace && owner)
            for (int r = tid; r < R; r += nt)
                I.trace[(size_t)rnd * R + r] = (double)E[rep[r]] * I.unit;
        __syncthreads();
        if (rnd < 0) {
            if (reached()) {
                rounds_run = 0;
                break;
            }
            continue;
        }

        // --- replica exchange and round-trip labels ---------------------------------------
        exchange_pairs(I, rnd, E, rep, rung, tried, accd);
        __syncthreads();
        update_labels(R, rung, label, trips, nullptr, up, down);
        __syncthreads();
        if (reached()) {
            rounds_run = rnd + 1;
            break;
        }
    }

    // --- results ---------------------------------------------------------------------------
    int64_t* out = I.out;
    if (owner) {
        if (tid == 0) {
            out[0] = sc->bestE; out[1] = sc->bround; out[2] = sc->brung; out[3] = sc->bslot;
            out[4] = sc->have;
            out[5] = rounds_run;
        }
        for (int x = tid; x < R; x += nt) {
            pt_out_arr(out, R, PT_OUT_E)[x] = E[x];
            pt_out_arr(out, R, PT_OUT_RUNG)[x] = rung[x];
            pt_out_arr(out, R, PT_OUT_TRIED)[x] = tried[x];
            pt_out_arr(out, R, PT_OUT_ACC)[x] = accd[x];
            pt_out_arr(out, R, PT_OUT_TRIPS)[x] = trips[x];
            pt_out_arr(out, R, PT_OUT_UP)[x] = up[x];
            pt_out_arr(out, R, PT_OUT_DOWN)[x] = down[x];
        }
    }
    if (I.final_s) {
        const int len = pp1 - pp0;
        for (int idx = tid; idx < len * rg; idx += nt) {
            const int s = idx / len, p = pp0 + idx - s * len;
            if (a0 + s < R) I.final_s[(size_t)(a0 + s) * n + p] = sp[(size_t)p * rg + s];
        }
    }
}

template <int WM, bool INIT>
__global__ __launch_bounds__(1024) void pt_lds_kernel(const PtInstance* __restrict__ insts,
                                                      const int32_t* __restrict__ order,
                                                      int n_inst, int32_t* queue) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ LShared sc;
    __shared__ long long scrE[16];
    __shared__ int scrR[16];
    const SplitArgs none{nullptr, 0ull};
    for (;;) {
        if (threadIdx.x == 0) sc.inst = atomicAdd(queue, 1);
        __syncthreads();
        const int qi = sc.inst;
        if (qi >= n_inst) return;
        const PtInstance I = insts[order[qi]];
        lds_instance<WM, INIT, false>(I, 0, smem, &sc, scrE, scrR, none);
        __syncthreads();
    }
}

// Split engine: work items are (instance, group), groups of one instance consecutive in the
// queue, so a workgroup waiting in a mailbox waits only on items already taken by resident
// workgroups or on the next ones taken (the grid holds every group of the largest instance).
template <int WM, bool INIT, bool CL, bool PARTS = false>
__global__ __launch_bounds__(1024) void pt_split_kernel(const PtInstance* __restrict__ insts,
                                                        const int32_t* __restrict__ items,
                                                        int n_items, int32_t* queue, SplitArgs SA) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ LShared sc;
    __shared__ long long scrE[16];
    __shared__ int scrR[16];
    for (;;) {
        if (threadIdx.x == 0) {
            int got = n_items;
            if (INIT || __hip_atomic_load(SA.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                const int home = (int)(blockIdx.x % (unsigned)SA.nq);
                for (int d = 0; d < SA.nq; ++d) {
                    const int q = home + d < SA.nq ? home + d : home + d - SA.nq;
                    const int len = SA.qoff[q + 1] - SA.qoff[q];
                    if (__hip_atomic_load(queue + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len)
                        continue;
                    const int i = atomicAdd(queue + q, 1);
                    if (i < len) {
                        got = SA.qoff[q] + i;
                        break;
                    }
                }
            }
            sc.inst = got;
        }
        __syncthreads();
        const int qi = sc.inst;
        if (qi >= n_items) return;
        const int item = items[qi];
        const PtInstance I = insts[item >> 10];
        lds_instance<WM, INIT, true, CL, PARTS>(I, item & 1023, smem, &sc, scrE, scrR, SA);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Global engine
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void g_init_spins(PtInstance I, GState G) {
    const int64_t total = (int64_t)I.n * I.Rp;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(idx / I.Rp), a = (int)(idx - (int64_t)p * I.Rp);
        G.sp[idx] = a < I.R ? init_spin(I, p, a) : (int8_t)0;
    }
}

// Reduce a per-thread energy change over the slots of the block, then one atomic per slot into
// the block's sub-counter (integer sums: the order of arrival does not matter).
__device__ __forceinline__ void block_slot_reduce(int64_t v, int a, int R, int Rp, long long* red,
                                                  int64_t* dacc, int par) {
    if (v != 0) atomicAdd((unsigned long long*)&red[a], (unsigned long long)v);
    __syncthreads();
    int64_t* dst = dacc + ((size_t)par * PT_DACC_SPLIT + blockIdx.x % PT_DACC_SPLIT) * Rp;
    for (int x = threadIdx.x; x < R; x += blockDim.x)
        if (red[x] != 0) atomicAdd((unsigned long long*)&dst[x], (unsigned long long)red[x]);
}

template <int WM>
__global__ __launch_bounds__(1024) void g_init_energy(PtInstance I, GState G) {
    __shared__ long long red[1024];
    const int Rp = I.Rp, R = I.R, tid = threadIdx.x;
    for (int x = tid; x < Rp; x += blockDim.x) red[x] = 0;
    __syncthreads();
    const int a = tid % Rp, qo = tid / Rp, qpb = blockDim.x / Rp;
    int64_t acc = 0;
    if (a < R)
        for (int p = blockIdx.x * qpb + qo; p < I.n; p += gridDim.x * qpb)
            acc += site_energy<WM, 8>(p, a, Rp, G.sp, I.row_ptr, I.edge, I.h);
    block_slot_reduce(acc, a, R, Rp, red, G.dacc, 1);   // parity 1 holds the init totals
}

// With a target: true once the round end of a round before rnd reached it; the launches of round
// rnd then return at entry, every workgroup alike (the flag was written by an earlier launch).
__device__ __forceinline__ bool g_stopped(const GState& G, int rnd) {
    if (!G.stop) return false;
    const int s = *G.stop;
    return s != 0 && rnd >= s - 1;
}

// Spin of slot a at position p in either layout.
__device__ __forceinline__ int8_t g_spin(const GState& G, int p, int Rp, int a) {
    return G.spw ? (((G.spw[p] >> a) & 1u) ? (int8_t)-1 : (int8_t)1) : G.sp[(size_t)p * Rp + a];
}

// Colour step of the CSR engine with one 32-bit word per position (R <= 32, Rp = 32): the 32 lanes
// of a position read each neighbour's word once (a broadcast), take their own bit, decide as
// g_color_step does, and combine their flips with a ballot; lane 0 of the position stores the
// word.  The whole spin array is n words, so the instance stays in each XCD's L2.  The first GW
// edge records of the row come from I.ell (row-major, zero-padded, neighbour as its word index),
// so the neighbour words are one dependent load away; row pointers are read, alongside, only when
// some row is longer than GW, and such rows finish from the CSR in GW-edge chunks.  A padding record
// adds +-0 to the field, which changes at most the sign of a zero field (accepted with zero energy
// change either way).
template <int WM, int GW>
__global__ __launch_bounds__(1024) void g_color_bits(PtInstance I, GState G, int cs, int ce, int j,
                                                     int rl) {
    __shared__ long long red[1024];
    const int rnd = *G.round_base + rl, cur = rnd & 1;
    if (g_stopped(G, rnd)) return;
    const uint32_t t = (uint32_t)(rnd * I.S + j);
    const int Rp = I.Rp, R = I.R, tid = threadIdx.x;
    using LT = typename std::conditional<WM == PT_WM_F32, double, int32_t>::type;   // field
    using WT = typename std::conditional<WM == PT_WM_F32, float, int32_t>::type;    // weight
    constexpr int EW = WM == PT_WM_I8 ? 1 : 2;   // 32-bit words per edge record
    constexpr int NW = GW * EW;
    const int a = tid & 31;
    const int p = cs + blockIdx.x * (int)(blockDim.x >> 5) + (tid >> 5);
    const bool live = p < ce;
    uint32_t own = 0;
    bool flip = false;
    int64_t d = 0;
    if (live) {
        uint32_t ev[NW];
        const pt_v4u* src = (const pt_v4u*)I.ell + (size_t)p * (NW / 4);
#pragma unroll
        for (int i = 0; i < NW / 4; ++i) {
            const pt_v4u x = src[i];
            ev[4 * i] = x.x;
            ev[4 * i + 1] = x.y;
            ev[4 * i + 2] = x.z;
            ev[4 * i + 3] = x.w;
        }
        own = G.spw[p];
        int t0 = 0, t1 = 0;
        if (I.ell_rest) {
            t0 = I.row_ptr[p];
            t1 = I.row_ptr[p + 1];
        }
        LT l = I.h ? (LT)((const WT*)I.h)[p] : (LT)0;
        uint32_t v[GW];
#pragma unroll
        for (int u = 0; u < GW; ++u) v[u] = G.spw[WM == PT_WM_I8 ? (ev[u] >> 8) : ev[EW * u]];
#pragma unroll
        for (int u = 0; u < GW; ++u) {
            const bool neg = (v[u] >> a) & 1u;
            if constexpr (WM == PT_WM_F32) {
                l += (double)__uint_as_float(ev[EW * u + 1] ^ ((uint32_t)neg << 31));
            } else {
                const int32_t w = WM == PT_WM_I8 ? (int32_t)(int8_t)(ev[u] & 0xffu)
                                                 : (int32_t)ev[EW * u + 1];
                l += neg ? -w : w;
            }
        }
        for (int k = t0 + GW; k < t1; k += GW) {
            int cc[GW];
            WT w[GW];
            uint32_t vv[GW];
#pragma unroll
            for (int u = 0; u < GW; ++u) {
                const bool ok = k + u < t1;
                if constexpr (WM == PT_WM_F32) cc[u] = edge_float(I.edge, ok ? k + u : 0, w[u]);
                else cc[u] = edge_int<WM>(I.edge, ok ? k + u : 0, w[u]);
                cc[u] = ok ? cc[u] : 0;
            }
#pragma unroll
            for (int u = 0; u < GW; ++u) vv[u] = G.spw[cc[u]];
#pragma unroll
            for (int u = 0; u < GW; ++u) {
                const bool neg = (vv[u] >> a) & 1u;
                if constexpr (WM == PT_WM_F32) {
                    const double nl = l + (double)__uint_as_float(__float_as_uint(w[u]) ^ ((uint32_t)neg << 31));
                    l = k + u < t1 ? nl : l;
                } else {
                    l += k + u < t1 ? (neg ? -w[u] : w[u]) : 0;
                }
            }
        }
        const int8_t si = ((own >> a) & 1u) ? (int8_t)-1 : (int8_t)1;
        if (a < R) {
            const int g = G.rung[cur * Rp + a];
            const pt_u4 rv = pt_philox((uint32_t)(p >> 2), (uint32_t)a, t, PT_TAG_METRO, I.k0, I.k1);
            if constexpr (WM == PT_WM_F32) {
                const double q = si > 0 ? l : -l;
                if (pt_accept_float_dev((float)I.beta[g], (float)q, pt_word(rv, p & 3))) {
                    flip = true;
                    d = pt_fix37d(q);
                }
            } else {
                const uint32_t* thr_row = I.thr + (size_t)g * I.TQ;
                const int32_t q = si > 0 ? l : -l;
                if (q <= 0 || (q <= I.TQ && pt_word(rv, p & 3) < thr_row[q - 1])) {
                    flip = true;
                    d = 2 * (int64_t)q;
                }
            }
        }
    }
    // the position's 32 decisions as one mask; lane 0 of the position writes the word
    const uint64_t m = __ballot(flip);
    if (live && a == 0) {
        const uint32_t fm = (uint32_t)(m >> (tid & 32));
        if (fm) G.spw[p] = own ^ fm;
    }
    // the slot sums' LDS is cleared here, not at entry, so no barrier precedes the loads
    for (int x = tid; x < Rp; x += blockDim.x) red[x] = 0;
    __syncthreads();
    block_slot_reduce(d, a, R, Rp, red, G.dacc, cur);
}

__global__ __launch_bounds__(256) void g_pack_bits(PtInstance I, GState G, int unpack) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < I.n; p += gridDim.x * blockDim.x) {
        int8_t* row = G.sp + (size_t)p * I.Rp;
        if (unpack) {
            const uint32_t w = G.spw[p];
            for (int a = 0; a < I.Rp; ++a)
                row[a] = a < I.R ? (((w >> a) & 1u) ? (int8_t)-1 : (int8_t)1) : (int8_t)0;
        } else {
            uint32_t w = 0;
            for (int a = 0; a < I.R; ++a) w |= (row[a] < 0 ? 1u : 0u) << a;
            G.spw[p] = w;
        }
    }
}

// Colour step of the CSR engine: one Metropolis attempt per lane, so a launch has (colour size x
// slots) independent lanes in flight to cover the dependent row-pointer -> edge -> neighbour-spin
// loads.  Lane (r, a) updates rank r of colour c in slot a with word (p & 3) of the Philox block
// (p >> 2, a, t): the same uniform the LDS engine draws for that attempt.
template <int WM, bool UNI, int GW>
__global__ __launch_bounds__(1024) void g_color_step(PtInstance I, GState G, int cs, int ce, int j,
                                                     int rl) {
    __shared__ long long red[1024];
    const int rnd = *G.round_base + rl, cur = rnd & 1;
    if (g_stopped(G, rnd)) return;
    const uint32_t t = (uint32_t)(rnd * I.S + j);
    const int Rp = I.Rp, R = I.R, tid = threadIdx.x;
    for (int x = tid; x < Rp; x += blockDim.x) red[x] = 0;
    __syncthreads();
    // with Rp a multiple of 64 every wave holds one position: make that visible to the compiler
    // so row pointers and edges go through the scalar unit
    const int a = tid % Rp;
    const int ro = UNI ? __builtin_amdgcn_readfirstlane(tid / Rp) : tid / Rp;
    const int p = cs + blockIdx.x * (blockDim.x / Rp) + ro;
    int64_t d = 0;
    if (a < R && p < ce) {
        const int g = G.rung[cur * Rp + a];
        const float bf = (float)I.beta[g];
        const uint32_t* thr_row = WM == PT_WM_F32 ? nullptr : I.thr + (size_t)g * I.TQ;
        const pt_u4 rv = pt_philox((uint32_t)(p >> 2), (uint32_t)a, t, PT_TAG_METRO, I.k0, I.k1);
        d = site_update_ell<WM, GW>(p, a, Rp, G.sp, I.row_ptr, I.edge, I.ell, I.ell_rest != 0, I.h,
                                    pt_word(rv, p & 3), thr_row, I.TQ, bf);
    }
    block_slot_reduce(d, a, R, Rp, red, G.dacc, cur);
}

// Colour step on a periodic hypercubic lattice (L even) with +/-1 couplings and no field.
// Neighbour positions follow from coordinates, so no column indices are read; a site's six bond
// signs are one byte.  Positions of colour c are the parity-c sites in index order, so rank r of
// colour c sits in row r / (L/2) at x = 2 (r mod L/2) + parity.  With 64 slots per position all
// coordinate arithmetic is wave-uniform and runs on the scalar unit.
template <int DIM, bool UNI>
__global__ __launch_bounds__(1024) void g_lattice_step(PtInstance I, GState G, int c, int j,
                                                       int rl) {
    __shared__ long long red[1024];
    const int rnd = *G.round_base + rl, cur = rnd & 1;
    if (g_stopped(G, rnd)) return;
    const uint32_t t = (uint32_t)(rnd * I.S + j);
    const int Rp = I.Rp, R = I.R, tid = threadIdx.x;
    for (int x = tid; x < Rp; x += blockDim.x) red[x] = 0;
    __syncthreads();
    const int a = tid % Rp, qpb = blockDim.x / Rp;
    const int qo = UNI ? __builtin_amdgcn_readfirstlane(tid / Rp) : tid / Rp;
    const int L = I.lL, half = L >> 1, nhalf = I.n >> 1;
    const int cs = c * nhalf, ce = cs + nhalf, ob = (1 - c) * nhalf;
    int64_t dsum = 0;
    if (a < R) {
        const int g = G.rung[cur * Rp + a];
        const uint32_t* tr = I.thr + (size_t)g * I.TQ;
        const uint32_t t2 = I.TQ >= 2 ? tr[1] : 0u;
        const uint32_t t4 = I.TQ >= 4 ? tr[3] : 0u;
        const uint32_t t6 = (DIM == 3 && I.TQ >= 6) ? tr[5] : 0u;
        int8_t* base = G.sp + a;
        const int q1 = (ce - 1) >> 2;
        for (int qq = (cs >> 2) + blockIdx.x * qpb + qo; qq <= q1; qq += gridDim.x * qpb) {
            const pt_u4 rv = pt_philox((uint32_t)qq, (uint32_t)a, t, PT_TAG_METRO, I.k0, I.k1);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int p = 4 * qq + kk;
                if (p < cs || p >= ce) continue;
                const int r = p - cs;
                const int row = r / half, jr = r - row * half;
                int x, y, z = 0;
                if (DIM == 3) {
                    y = row % L;
                    z = row / L;
                    x = 2 * jr + ((c + y + z) & 1);
                } else {
                    y = row;
                    x = 2 * jr + ((c + y) & 1);
                }
                const int plane = DIM == 3 ? z * L : 0;
                const uint32_t jb = I.jbits[(plane + y) * L + x];
                const int xp = x + 1 == L ? 0 : x + 1, xm = x == 0 ? L - 1 : x - 1;
                const int yp = y + 1 == L ? 0 : y + 1, ym = y == 0 ? L - 1 : y - 1;
                const int rowpos = ob + (plane + y) * half;
                int n0 = base[(size_t)(rowpos + (xp >> 1)) * Rp];
                int n1 = base[(size_t)(rowpos + (xm >> 1)) * Rp];
                int n2 = base[(size_t)(ob + (plane + yp) * half + (x >> 1)) * Rp];
                int n3 = base[(size_t)(ob + (plane + ym) * half + (x >> 1)) * Rp];
                int l = ((jb & 1u) ? -n0 : n0) + ((jb & 2u) ? -n1 : n1) +
                        ((jb & 4u) ? -n2 : n2) + ((jb & 8u) ? -n3 : n3);
                if (DIM == 3) {
                    const int zp = z + 1 == L ? 0 : z + 1, zm = z == 0 ? L - 1 : z - 1;
                    const int n4 = base[(size_t)(ob + (zp * L + y) * half + (x >> 1)) * Rp];
                    const int n5 = base[(size_t)(ob + (zm * L + y) * half + (x >> 1)) * Rp];
                    l += ((jb & 16u) ? -n4 : n4) + ((jb & 32u) ? -n5 : n5);
                }
                const int8_t si = base[(size_t)p * Rp];
                const int q = si > 0 ? l : -l;
                const uint32_t th = q == 2 ? t2 : (q == 4 ? t4 : t6);
                if (q <= 0 || pt_word(rv, kk) < th) {
                    base[(size_t)p * Rp] = (int8_t)-si;
                    dsum += 2 * q;
                }
            }
        }
    }
    block_slot_reduce(dsum, a, R, Rp, red, G.dacc, cur);
}

// Round end (rnd >= 0) or initialisation end (rnd == -1).  Every workgroup evaluates the energies
// and the best-rung scan (a few hundred bytes) and copies its share of the best configuration if
// it improved; workgroup 0 alone writes the next round's state, so no workgroup reads what another
// one of the same launch writes.
template <int WM>
__global__ __launch_bounds__(1024) void g_round_end(PtInstance I, GState G, int rl, int init_end) {
    __shared__ int64_t Es[1024];
    __shared__ int32_t reps[1024];
    __shared__ int32_t rungs[1024];
    __shared__ long long scrE[16];
    __shared__ int scrR[16];
    const int R = I.R, Rp = I.Rp, tid = threadIdx.x;
    const bool init = init_end != 0;
    const int rnd = init ? -1 : *G.round_base + rl;
    if (!init && g_stopped(G, rnd)) return;
    const int cur = rnd & 1, nxt = cur ^ 1;
    if (R <= PT_WAVE) {
        // wave 0 evaluates the round end (WaveLadder) with the ladder loaded from this round's
        // buffers; workgroup 0 writes the next round's; every workgroup copies its share of an
        // improved best configuration
        __shared__ int sh_imp, sh_slot;
        int64_t* out = I.out;
        if (tid < PT_WAVE) {
            const int r = tid;
            const bool in = r < R;
            WaveLadder W;
            W.rep = init || !in ? r : G.rep[(size_t)cur * Rp + r];
            int64_t E = 0;
            if (in) {
                E = init ? 0 : G.Eb[(size_t)cur * Rp + W.rep];
                int64_t d = 0;
#pragma unroll 8
                for (int k = 0; k < PT_DACC_SPLIT; ++k) d += G.dacc[((size_t)cur * PT_DACC_SPLIT + k) * Rp + W.rep];
                E += d;
                if (init && WM != PT_WM_F32) E = -(E / 2);
            }
            W.Eb = E;
            W.beta = in ? I.beta[r] : 0.0;
            W.lab = init ? ((R > 1 && r == 0) ? 1 : ((R > 1 && r == R - 1) ? -1 : 0))
                         : (in ? G.label[W.rep] : 0);
            W.trips = in ? (int32_t)pt_out_arr(out, R, PT_OUT_TRIPS)[W.rep] : 0;
            W.tried = in ? (int32_t)pt_out_arr(out, R, PT_OUT_TRIED)[r] : 0;
            W.acc = in ? (int32_t)pt_out_arr(out, R, PT_OUT_ACC)[r] : 0;
            W.up = in ? (int32_t)pt_out_arr(out, R, PT_OUT_UP)[r] : 0;
            W.down = in ? (int32_t)pt_out_arr(out, R, PT_OUT_DOWN)[r] : 0;
            const int64_t* bc = G.best + 5 * cur;
            W.have = init ? 0 : (int32_t)bc[4];
            W.bestE = bc[0];
            W.best_round = (int32_t)bc[1];
            W.best_rung = (int32_t)bc[2];
            W.best_slot = (int32_t)bc[3];
            const bool imp = wave_ladder_round(W, I, rnd + 1, E, r, blockIdx.x == 0);
            if (r == 0) {
                sh_imp = imp;
                sh_slot = W.best_slot;
            }
            if (blockIdx.x == 0) {
                if (in) {
                    G.Eb[(size_t)nxt * Rp + W.rep] = W.Eb;
                    G.rung[(size_t)nxt * Rp + W.rep] = r;
                    G.rep[(size_t)nxt * Rp + r] = W.rep;
                    G.label[W.rep] = W.lab;
                    pt_out_arr(out, R, PT_OUT_TRIPS)[W.rep] = W.trips;
                    pt_out_arr(out, R, PT_OUT_TRIED)[r] = W.tried;
                    pt_out_arr(out, R, PT_OUT_ACC)[r] = W.acc;
                    pt_out_arr(out, R, PT_OUT_UP)[r] = W.up;
                    pt_out_arr(out, R, PT_OUT_DOWN)[r] = W.down;
                    for (int k = 0; k < PT_DACC_SPLIT; ++k)
                        G.dacc[((size_t)nxt * PT_DACC_SPLIT + k) * Rp + r] = 0;
                }
                if (r == 0) {
                    int64_t* bn = G.best + 5 * nxt;
                    bn[0] = W.bestE; bn[1] = W.best_round; bn[2] = W.best_rung; bn[3] = W.best_slot;
                    bn[4] = W.have;
                    if (G.stop && I.has_target && W.have && W.bestE <= I.target_units) *G.stop = rnd + 2;
                }
            }
        }
        __syncthreads();
        if (sh_imp) {
            const int slot = sh_slot;
            const int chunk = (I.n + gridDim.x - 1) / gridDim.x;
            const int p0 = blockIdx.x * chunk, p1 = min(I.n, p0 + chunk);
            for (int p = p0 + tid; p < p1; p += blockDim.x) I.best_s[p] = g_spin(G, p, Rp, slot);
        }
        return;
    }
    for (int x = tid; x < R; x += blockDim.x) {
        Es[x] = init ? 0 : G.Eb[(size_t)cur * Rp + x];
        reps[x] = init ? x : G.rep[(size_t)cur * Rp + x];
        rungs[x] = init ? x : G.rung[(size_t)cur * Rp + x];
    }
    __syncthreads();
    // the round's energy changes: one load per (counter, slot) pair, summed in LDS
    for (int idx = tid; idx < PT_DACC_SPLIT * R; idx += blockDim.x) {
        const int sp = idx / R, x = idx - sp * R;
        const int64_t v = G.dacc[((size_t)cur * PT_DACC_SPLIT + sp) * Rp + x];
        if (v != 0) atomicAdd((unsigned long long*)&Es[x], (unsigned long long)v);
    }
    __syncthreads();
    if (init && WM != PT_WM_F32) {
        for (int x = tid; x < R; x += blockDim.x) Es[x] = -(Es[x] / 2);
        __syncthreads();
    }
    int64_t mE;
    int32_t mR;
    block_best_rung(Es, reps, R, scrE, scrR, mE, mR);
    const int64_t* bc = G.best + 5 * cur;
    const bool have = !init && bc[4] != 0;
    const bool improved = !have || mE < bc[0];
    const int slot = reps[mR];
    if (improved) {
        const int chunk = (I.n + gridDim.x - 1) / gridDim.x;
        const int p0 = blockIdx.x * chunk, p1 = min(I.n, p0 + chunk);
        for (int p = p0 + tid; p < p1; p += blockDim.x)
            I.best_s[p] = g_spin(G, p, Rp, slot);
    }
    if (blockIdx.x != 0) return;

    int64_t* out = I.out;
    if (tid == 0) {
        int64_t* bn = G.best + 5 * nxt;
        if (improved) {
            bn[0] = mE; bn[1] = rnd + 1; bn[2] = mR; bn[3] = slot; bn[4] = 1;
        } else {
            for (int x = 0; x < 5; ++x) bn[x] = bc[x];
        }
        if (G.stop && I.has_target && (improved ? mE : bc[0]) <= I.target_units) *G.stop = rnd + 2;
        if (!init && I.hist) I.hist[rnd] = (double)(improved ? mE : bc[0]) * I.unit;
    }
    if (!init && I.trace)
        for (int r = tid; r < R; r += blockDim.x)
            I.trace[(size_t)rnd * R + r] = (double)Es[reps[r]] * I.unit;
    if (init) {
        for (int x = tid; x < R; x += blockDim.x)
            G.label[x] = (R > 1 && x == 0) ? 1 : ((R > 1 && x == R - 1) ? -1 : 0);
    } else {
        __syncthreads();   // the trace above reads reps before the exchange rewrites them
        exchange_pairs(I, rnd, Es, reps, rungs, pt_out_arr(out, R, PT_OUT_TRIED),
                       pt_out_arr(out, R, PT_OUT_ACC));
        __syncthreads();
        update_labels(R, rungs, G.label, nullptr, pt_out_arr(out, R, PT_OUT_TRIPS),
                      pt_out_arr(out, R, PT_OUT_UP), pt_out_arr(out, R, PT_OUT_DOWN));
        __syncthreads();
    }
    for (int x = tid; x < R; x += blockDim.x) {
        G.Eb[(size_t)nxt * Rp + x] = Es[x];
        for (int k = 0; k < PT_DACC_SPLIT; ++k) G.dacc[((size_t)nxt * PT_DACC_SPLIT + k) * Rp + x] = 0;
        G.rung[(size_t)nxt * Rp + x] = rungs[x];
        G.rep[(size_t)nxt * Rp + x] = reps[x];
    }
}

__global__ void g_tick(int32_t* round_base, int rounds) {
    if (threadIdx.x == 0) *round_base += rounds;
}

// Round-end workgroup: at least 256 threads (the best-state copy is spread over them), one per
// slot beyond that.
inline int block_for_slots(int Rp) { return Rp <= 256 ? 256 : ((Rp + 63) / 64) * 64; }

}  // namespace

size_t pt_lds_bytes(int n, int Rp, int k) {
    const size_t b = (((size_t)n * Rp + 15) & ~(size_t)15) + (size_t)Rp * (5 * 8 + 4 * 4 + 4);
    return ((b + 15) & ~(size_t)15) + (((size_t)4 * (k + 1) + 15) & ~(size_t)15);
}

size_t pt_split_lds_bytes(int n, int rg, int Rp, int k, int thr_words, bool cl, int nnz,
                          bool has_h, bool parts) {
    size_t b = pt_lds_img_bytes(n, rg) + (size_t)Rp * (5 * 8 + 4 * 4 + 4);
    b = (b + 15) & ~(size_t)15;
    b += ((size_t)4 * (k + 1) + 15) & ~(size_t)15;
    b += ((size_t)4 * thr_words + 15) & ~(size_t)15;
    if (parts) b += ((size_t)8 * (Rp + 3 * rg) + 15) & ~(size_t)15;
    if (cl) {
        b += ((size_t)4 * (n + 1) + 15) & ~(size_t)15;
        b += ((size_t)2 * nnz + 15) & ~(size_t)15;
        if (has_h) b += (size_t)4 * n;
    }
    return b;
}

template <int WM, bool INIT, bool CL, bool PARTS = false>
static hipError_t split_launch(const PtInstance* d_insts, const int32_t* d_items, int n_items,
                               int32_t* d_queue, int block, int grid, size_t lds_bytes,
                               const SplitArgs& A, bool cooperative, hipStream_t s) {
    const void* fn = (const void*)pt_split_kernel<WM, INIT, CL, PARTS>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    if (cooperative) {
        SplitArgs a = A;
        void* args[] = {(void*)&d_insts, (void*)&d_items, (void*)&n_items, (void*)&d_queue, (void*)&a};
        return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, (unsigned)lds_bytes, s);
    }
    pt_split_kernel<WM, INIT, CL, PARTS><<<grid, block, lds_bytes, s>>>(d_insts, d_items, n_items, d_queue, A);
    return hipGetLastError();
}

hipError_t pt_launch_split(const PtInstance* d_insts, const int32_t* d_items, int n_items,
                           int32_t* d_queue, int wm, int block, int grid, size_t lds_bytes,
                           bool init, const SplitArgs& A, bool cooperative, hipStream_t s,
                           bool cl, bool parts) {
    if (init) cooperative = false;   // no group waits on another while drawing initial states
    if (parts && !init) {   // the init launch draws by group, as without parts
        if (cl) return hipErrorInvalidValue;
        if (wm == PT_WM_I8)
            return split_launch<PT_WM_I8, false, false, true>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
        if (wm == PT_WM_I32)
            return split_launch<PT_WM_I32, false, false, true>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
        return split_launch<PT_WM_F32, false, false, true>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
    }
    if (cl) {   // couplings in LDS: int8 couplings only; the init launch draws from the CSR in HBM
        if (wm != PT_WM_I8) return hipErrorInvalidValue;
        return init ? split_launch<PT_WM_I8, true, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s)
                    : split_launch<PT_WM_I8, false, true>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
    }
    if (wm == PT_WM_I8)
        return init ? split_launch<PT_WM_I8, true, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s)
                    : split_launch<PT_WM_I8, false, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
    if (wm == PT_WM_I32)
        return init ? split_launch<PT_WM_I32, true, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s)
                    : split_launch<PT_WM_I32, false, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
    return init ? split_launch<PT_WM_F32, true, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s)
                : split_launch<PT_WM_F32, false, false>(d_insts, d_items, n_items, d_queue, block, grid, lds_bytes, A, cooperative, s);
}

hipError_t pt_split_occupancy(int wm, int block, size_t lds_bytes, int* per_cu, bool cl,
                              bool parts) {
    const void* fn = parts             ? (wm == PT_WM_I8    ? (const void*)pt_split_kernel<PT_WM_I8, false, false, true>
                                          : wm == PT_WM_I32 ? (const void*)pt_split_kernel<PT_WM_I32, false, false, true>
                                                            : (const void*)pt_split_kernel<PT_WM_F32, false, false, true>)
                     : cl              ? (const void*)pt_split_kernel<PT_WM_I8, false, true>
                     : wm == PT_WM_I8  ? (const void*)pt_split_kernel<PT_WM_I8, false, false>
                     : wm == PT_WM_I32 ? (const void*)pt_split_kernel<PT_WM_I32, false, false>
                                       : (const void*)pt_split_kernel<PT_WM_F32, false, false>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, block, lds_bytes);
}

template <int WM, bool INIT>
static hipError_t lds_launch(const PtInstance* d_insts, const int32_t* d_order, int n_inst,
                             int32_t* d_queue, int block, int grid, size_t lds_bytes,
                             hipStream_t s) {
    hipError_t e = hipFuncSetAttribute((const void*)pt_lds_kernel<WM, INIT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    pt_lds_kernel<WM, INIT><<<grid, block, lds_bytes, s>>>(d_insts, d_order, n_inst, d_queue);
    return hipGetLastError();
}

hipError_t pt_launch_lds(const PtInstance* d_insts, const int32_t* d_order, int n_inst,
                         int32_t* d_queue, int wm, int block, int grid, size_t lds_bytes,
                         bool init, hipStream_t s) {
    if (wm == PT_WM_I8)
        return init ? lds_launch<PT_WM_I8, true>(d_insts, d_order, n_inst, d_queue, block, grid, lds_bytes, s)
                    : lds_launch<PT_WM_I8, false>(d_insts, d_order, n_inst, d_queue, block, grid, lds_bytes, s);
    if (wm == PT_WM_I32)
        return init ? lds_launch<PT_WM_I32, true>(d_insts, d_order, n_inst, d_queue, block, grid, lds_bytes, s)
                    : lds_launch<PT_WM_I32, false>(d_insts, d_order, n_inst, d_queue, block, grid, lds_bytes, s);
    return init ? lds_launch<PT_WM_F32, true>(d_insts, d_order, n_inst, d_queue, block, grid, lds_bytes, s)
                : lds_launch<PT_WM_F32, false>(d_insts, d_order, n_inst, d_queue, block, grid, lds_bytes, s);
}

template <int WM>
static hipError_t g_init_impl(const PtInstance& I, const GState& G, int block, int reg,
                              hipStream_t s) {
    const int64_t total = (int64_t)I.n * I.Rp;
    const int gs = (int)std::min<int64_t>((total + 255) / 256, 4096);
    g_init_spins<<<gs, 256, 0, s>>>(I, G);
    const int qpb = block / I.Rp;
    const int ge = (int)std::min<int64_t>((I.n + qpb - 1) / qpb, 2048);
    g_init_energy<WM><<<ge, block, 0, s>>>(I, G);
    g_round_end<WM><<<reg, block_for_slots(I.Rp), 0, s>>>(I, G, 0, 1);
    return hipGetLastError();
}

hipError_t pt_launch_g_init(const PtInstance& I, const GState& G, int block, int reg,
                            hipStream_t s) {
    if (I.wm == PT_WM_I8) return g_init_impl<PT_WM_I8>(I, G, block, reg, s);
    if (I.wm == PT_WM_I32) return g_init_impl<PT_WM_I32>(I, G, block, reg, s);
    return g_init_impl<PT_WM_F32>(I, G, block, reg, s);
}

template <bool UNI>
static void g_color_dispatch(const PtInstance& I, const GState& G, int cs, int ce, int j, int rl,
                             int block, int grid, hipStream_t s) {
    if (I.gq == 4) {
        if (I.wm == PT_WM_I8) g_color_step<PT_WM_I8, UNI, 4><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
        else if (I.wm == PT_WM_I32) g_color_step<PT_WM_I32, UNI, 4><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
        else g_color_step<PT_WM_F32, UNI, 4><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
    } else {
        if (I.wm == PT_WM_I8) g_color_step<PT_WM_I8, UNI, 8><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
        else if (I.wm == PT_WM_I32) g_color_step<PT_WM_I32, UNI, 8><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
        else g_color_step<PT_WM_F32, UNI, 8><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
    }
}

hipError_t pt_launch_g_pack(const PtInstance& I, const GState& G, bool unpack, hipStream_t s) {
    const int grid = std::max(1, std::min(1024, (I.n + 255) / 256));
    g_pack_bits<<<grid, 256, 0, s>>>(I, G, unpack ? 1 : 0);
    return hipGetLastError();
}

hipError_t pt_launch_g_color(const PtInstance& I, const GState& G, int cs, int ce, int j, int rl,
                             int block, int grid, hipStream_t s) {
    if (G.spw) {
        if (I.gq == 4) {
            if (I.wm == PT_WM_I8) g_color_bits<PT_WM_I8, 4><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
            else if (I.wm == PT_WM_I32) g_color_bits<PT_WM_I32, 4><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
            else g_color_bits<PT_WM_F32, 4><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
        } else {
            if (I.wm == PT_WM_I8) g_color_bits<PT_WM_I8, 8><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
            else if (I.wm == PT_WM_I32) g_color_bits<PT_WM_I32, 8><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
            else g_color_bits<PT_WM_F32, 8><<<grid, block, 0, s>>>(I, G, cs, ce, j, rl);
        }
        return hipGetLastError();
    }
    if (I.Rp % 64 == 0) g_color_dispatch<true>(I, G, cs, ce, j, rl, block, grid, s);
    else g_color_dispatch<false>(I, G, cs, ce, j, rl, block, grid, s);
    return hipGetLastError();
}

hipError_t pt_launch_g_lattice(const PtInstance& I, const GState& G, int c, int j, int rl,
                               int block, int grid, hipStream_t s) {
    const bool uni = I.Rp % 64 == 0;
    if (I.ldim == 3) {
        if (uni) g_lattice_step<3, true><<<grid, block, 0, s>>>(I, G, c, j, rl);
        else g_lattice_step<3, false><<<grid, block, 0, s>>>(I, G, c, j, rl);
    } else {
        if (uni) g_lattice_step<2, true><<<grid, block, 0, s>>>(I, G, c, j, rl);
        else g_lattice_step<2, false><<<grid, block, 0, s>>>(I, G, c, j, rl);
    }
    return hipGetLastError();
}

hipError_t pt_launch_g_round_end(const PtInstance& I, const GState& G, int rl,
                                 int round_end_grid, hipStream_t s) {
    const int b = block_for_slots(I.Rp);
    if (I.wm == PT_WM_I8) g_round_end<PT_WM_I8><<<round_end_grid, b, 0, s>>>(I, G, rl, 0);
    else if (I.wm == PT_WM_I32) g_round_end<PT_WM_I32><<<round_end_grid, b, 0, s>>>(I, G, rl, 0);
    else g_round_end<PT_WM_F32><<<round_end_grid, b, 0, s>>>(I, G, rl, 0);
    return hipGetLastError();
}

hipError_t pt_launch_g_tick(const GState& G, int rounds, hipStream_t s) {
    g_tick<<<1, 64, 0, s>>>(G.round_base, rounds);
    return hipGetLastError();
}
