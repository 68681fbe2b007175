// pt_coop.hip -- persistent CSR engine for instances whose replicas do not fit one workgroup.
//
// One cooperative launch runs every round.  Workgroup w owns a fixed slice of every colour class
// for the whole run.  Its LDS holds the row offsets and edge records of its slice, the spins of
// its own positions (all slots), the Metropolis thresholds and the ladder, so a colour step reads
// only neighbour spins from memory.  Spins also live in HBM as [position][slot] (Rp bytes per
// position): the owner writes every flip through to it for the other workgroups.  Colour steps
// are separated by a grid barrier; the barrier after the last colour of a round also publishes
// the round's energy sums, and then every workgroup evaluates the same round end (best state,
// exchange, labels) from those sums -- the decisions are deterministic, so the copies agree --
// and workgroup 0 alone keeps the statistics.  Lane (i, a) updates rank i of its slice in slot a
// with word (p & 3) of the Philox block (p >> 2, a, t), the uniform every engine draws for that
// attempt, so results are bit-identical to the other engines.
//
// Reference counterpart: the per-launch sweep and swap kernels of cuda/ising_pt.cu:86-146.
#include "pt_engine.h"
#include "pt_rules.h"
#include "pt_device.h"

#include <type_traits>

namespace {

using namespace pt_dev;

struct CoopShared {
    long long best[5];   // E, round, rung, slot, have
    int abort;
    int imp, slot;       // the round end's best-rung result, for the whole workgroup
};

// First rank and size of workgroup w's slice of colour class [cs, cs + cz).
__device__ __forceinline__ void slice(int cs, int cz, int w, int W, int& lo, int& cnt) {
    lo = cs + (int)((int64_t)cz * w / W);
    cnt = cs + (int)((int64_t)cz * (w + 1) / W) - lo;
}

template <int WM>
using field_t = std::conditional_t<WM == PT_WM_F32, float, int32_t>;    // one coupling
template <int WM>
using acc_t = std::conditional_t<WM == PT_WM_F32, double, int32_t>;     // a local field

template <int WM>
__device__ __forceinline__ int edge_at(const uint32_t* e, int t, field_t<WM>& w) {
    if constexpr (WM == PT_WM_F32) return edge_float(e, t, w);
    else return edge_int<WM>(e, t, w);
}

// Local fields of two rows [s0, e0) and [s1, e1) (an empty row gives its start value).  The first
// PT_PAIR_CHUNK edges of both rows are gathered together without branches -- edge records from
// LDS, then every neighbour spin from memory; a slot past the end of its row reads edge 0 and
// position 0 and is dropped by a select -- so all the loads issue back to back and the pair waits
// on memory once.  Longer rows finish edge by edge.  Terms are added in row order.
#define PT_PAIR_CHUNK 8

template <int WM>
__device__ __forceinline__ void fields2(int a, int Rp, const int8_t* sp, const uint32_t* edge,
                                        int s0, int e0, int s1, int e1, acc_t<WM>& l0,
                                        acc_t<WM>& l1) {
    constexpr int K = PT_PAIR_CHUNK;
    int c0[K], c1[K];
    int8_t v0[K], v1[K];
    field_t<WM> w0[K], w1[K];
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const bool ok0 = s0 + u < e0, ok1 = s1 + u < e1;
        c0[u] = edge_at<WM>(edge, ok0 ? s0 + u : 0, w0[u]);
        c1[u] = edge_at<WM>(edge, ok1 ? s1 + u : 0, w1[u]);
        c0[u] = ok0 ? c0[u] : 0;
        c1[u] = ok1 ? c1[u] : 0;
    }
#pragma unroll
    for (int u = 0; u < K; ++u) {
        v0[u] = ld_spin<true>(sp + (size_t)c0[u] * Rp + a);
        v1[u] = ld_spin<true>(sp + (size_t)c1[u] * Rp + a);
    }
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const acc_t<WM> n0 = v0[u] > 0 ? l0 + (acc_t<WM>)w0[u] : l0 - (acc_t<WM>)w0[u];
        const acc_t<WM> n1 = v1[u] > 0 ? l1 + (acc_t<WM>)w1[u] : l1 - (acc_t<WM>)w1[u];
        l0 = s0 + u < e0 ? n0 : l0;
        l1 = s1 + u < e1 ? n1 : l1;
    }
    for (int t = s0 + K; t < e0; ++t) {
        field_t<WM> w;
        const int c = edge_at<WM>(edge, t, w);
        l0 = ld_spin<true>(sp + (size_t)c * Rp + a) > 0 ? l0 + (acc_t<WM>)w : l0 - (acc_t<WM>)w;
    }
    for (int t = s1 + K; t < e1; ++t) {
        field_t<WM> w;
        const int c = edge_at<WM>(edge, t, w);
        l1 = ld_spin<true>(sp + (size_t)c * Rp + a) > 0 ? l1 + (acc_t<WM>)w : l1 - (acc_t<WM>)w;
    }
}

// Metropolis decision for spin si with local field l: true when the flip is accepted (a flip
// with zero energy change is accepted), with its energy change in units in d.
template <int WM>
__device__ __forceinline__ bool decide(int8_t si, acc_t<WM> l, uint32_t u,
                                       const uint32_t* thr_row, int TQ, float bf, int64_t& d) {
    if constexpr (WM == PT_WM_F32) {
        const double q = si > 0 ? l : -l;
        d = pt_fix37d(q);
        return pt_accept_float_dev(bf, (float)q, u);
    } else {
        const int32_t q = si > 0 ? l : -l;
        d = 2 * (int64_t)q;
        return q <= 0 || (q <= TQ && u < thr_row[q - 1]);
    }
}

template <int WM>
__device__ __forceinline__ acc_t<WM> field_h(const void* h, int p) {
    if constexpr (WM == PT_WM_F32) return h ? (double)((const float*)h)[p] : 0.0;
    else return h ? ((const int32_t*)h)[p] : 0;
}

template <int WM, bool UNI>
__global__ __launch_bounds__(1024) void coop_run(PtInstance I, GState G, CoopArgs C) {
    extern __shared__ __align__(16) char smem[];
    int32_t* cptr = (int32_t*)(smem + C.off_cptr);
    int32_t* lrp = (int32_t*)(smem + C.off_lrp);
    uint32_t* ledge = (uint32_t*)(smem + C.off_edge);
    int8_t* lsp = (int8_t*)(smem + C.off_lsp);
    uint32_t* thrL = (uint32_t*)(smem + C.off_thr);
    double* betaL = (double*)(smem + C.off_beta);
    float* betafL = (float*)(smem + C.off_betaf);
    int64_t* EbL = (int64_t*)(smem + C.off_Eb);
    int64_t* Es = (int64_t*)(smem + C.off_Es);
    long long* red = (long long*)(smem + C.off_red);
    int32_t* repL = (int32_t*)(smem + C.off_rep);
    int32_t* rungL = (int32_t*)(smem + C.off_rung);
    int32_t* labelL = (int32_t*)(smem + C.off_label);
    int64_t* stats = (int64_t*)(smem + C.off_stats);   // tried, acc, trips, up, down (Rp each)
    long long* scrE = (long long*)(smem + C.off_scr);
    int* scrR = (int*)(smem + C.off_scr + 16 * sizeof(long long));
    CoopShared* sh = (CoopShared*)(smem + C.off_sh);

    const int tid = threadIdx.x, nt = blockDim.x, w = blockIdx.x, W = gridDim.x;
    const int k = I.k, R = I.R, Rp = I.Rp;
    constexpr int ew = WM == PT_WM_I8 ? 1 : 2;   // 32-bit words per edge record
    const bool owner = w == 0;

    // ---- stage this workgroup's rows and spins, the thresholds, the ladder, round-0 state ----
    for (int x = tid; x <= k; x += nt) cptr[x] = I.color_ptr[x];
    __syncthreads();
    {
        int ob = 0, ebase = 0;
        for (int c = 0; c < k; ++c) {
            int lo, cnt;
            slice(cptr[c], cptr[c + 1] - cptr[c], w, W, lo, cnt);
            const int e0 = I.row_ptr[lo], e1 = I.row_ptr[lo + cnt];
            for (int i = tid; i <= cnt; i += nt) lrp[ob + c + i] = ebase + (I.row_ptr[lo + i] - e0);
            const uint32_t* src = (const uint32_t*)I.edge + (size_t)e0 * ew;
            uint32_t* dst = ledge + (size_t)ebase * ew;
            for (int i = tid; i < (e1 - e0) * ew; i += nt) dst[i] = src[i];
            for (int i = tid; i < cnt * Rp; i += nt) lsp[(size_t)ob * Rp + i] = G.sp[(size_t)lo * Rp + i];
            ob += cnt;
            ebase += e1 - e0;
        }
    }
    const uint32_t* thr = I.thr;
    if (WM != PT_WM_F32 && C.thr_lds) {
        for (int x = tid; x < R * I.TQ; x += nt) thrL[x] = I.thr[x];
        thr = thrL;
    }
    for (int x = tid; x < Rp; x += nt) {
        betaL[x] = x < R ? I.beta[x] : 0.0;
        betafL[x] = x < R ? (float)I.beta[x] : 0.0f;
        EbL[x] = G.Eb[x];
        repL[x] = G.rep[x];
        rungL[x] = G.rung[x];
        labelL[x] = G.label[x];
        red[x] = 0;
    }
    for (int x = tid; x < 5 * Rp; x += nt) stats[x] = 0;
    if (tid < 5) sh->best[tid] = G.best[tid];
    if (tid == 0) sh->abort = 0;
    __syncthreads();

    const int a = tid % Rp;
    const int pl = UNI ? __builtin_amdgcn_readfirstlane(tid / Rp) : tid / Rp;
    const int ppb = nt / Rp;
    const bool live = a < R;
    const int split = w % PT_DACC_SPLIT;
    uint32_t phase = 0;
    int64_t dsum = 0;
    for (int rnd = 0; rnd < I.n_rounds; ++rnd) {
        const int buf = rnd % 3;
        const int g = live ? rungL[a] : 0;
        const float bf = betafL[g];
        const uint32_t* thr_row = WM == PT_WM_F32 ? nullptr : thr + (size_t)g * I.TQ;
        for (int j = 0; j < I.S; ++j) {
            const uint32_t t = (uint32_t)(rnd * I.S + j);
            int ob = 0;
            for (int c = 0; c < k; ++c) {
                int lo, cnt;
                slice(cptr[c], cptr[c + 1] - cptr[c], w, W, lo, cnt);
                const int* rows = lrp + ob + c;
                int8_t* own = lsp + (size_t)ob * Rp + a;
                if (live)
                    for (int i0 = pl; i0 < cnt; i0 += 2 * ppb) {
                        const int i1 = i0 + ppb;
                        const bool has1 = i1 < cnt;
                        const int p0 = lo + i0, p1 = lo + i1;
                        const int s0 = rows[i0], e0 = rows[i0 + 1];
                        const int s1 = has1 ? rows[i1] : 0, e1 = has1 ? rows[i1 + 1] : 0;
                        acc_t<WM> l0 = field_h<WM>(I.h, p0);
                        acc_t<WM> l1 = has1 ? field_h<WM>(I.h, p1) : acc_t<WM>(0);
                        fields2<WM>(a, Rp, G.sp, ledge, s0, e0, s1, e1, l0, l1);
                        const pt_u4 r0 = pt_philox((uint32_t)(p0 >> 2), (uint32_t)a, t,
                                                   PT_TAG_METRO, I.k0, I.k1);
                        const int8_t si0 = own[(size_t)i0 * Rp];
                        int64_t d0;
                        if (decide<WM>(si0, l0, pt_word(r0, p0 & 3), thr_row, I.TQ, bf, d0)) {
                            own[(size_t)i0 * Rp] = (int8_t)-si0;
                            st_spin<true>(G.sp + (size_t)p0 * Rp + a, (int8_t)-si0);
                            dsum += d0;
                        }
                        if (has1) {
                            const pt_u4 r1 = pt_philox((uint32_t)(p1 >> 2), (uint32_t)a, t,
                                                       PT_TAG_METRO, I.k0, I.k1);
                            const int8_t si1 = own[(size_t)i1 * Rp];
                            int64_t d1;
                            if (decide<WM>(si1, l1, pt_word(r1, p1 & 3), thr_row, I.TQ, bf, d1)) {
                                own[(size_t)i1 * Rp] = (int8_t)-si1;
                                st_spin<true>(G.sp + (size_t)p1 * Rp + a, (int8_t)-si1);
                                dsum += d1;
                            }
                        }
                    }
                ob += cnt;
                if (j == I.S - 1 && c == k - 1) {
                    // the round's energy change of this workgroup, one atomic per slot
                    if (dsum != 0) atomicAdd((unsigned long long*)&red[a], (unsigned long long)dsum);
                    dsum = 0;
                    __syncthreads();
                    int64_t* dst = G.dacc + ((size_t)buf * PT_DACC_SPLIT + split) * Rp;
                    for (int x = tid; x < R; x += nt) {
                        if (red[x] != 0) atomicAdd((unsigned long long*)&dst[x], (unsigned long long)red[x]);
                        red[x] = 0;
                    }
                }
                if (!coop_sync(C, phase, &sh->abort)) return;
            }
        }

        // ---- round end: identical in every workgroup ---------------------------------------
        if (R <= PT_WAVE) {
            // by wave 0 (WaveLadder), the ladder kept in the LDS arrays between round ends
            if (tid < PT_WAVE) {
                const int r = tid;
                const bool in = r < R;
                WaveLadder Wl;
                Wl.rep = in ? repL[r] : r;
                int64_t E = 0;
                if (in) {
                    E = EbL[Wl.rep];
                    int64_t d = 0;
#pragma unroll 8
                    for (int sp = 0; sp < PT_DACC_SPLIT; ++sp)
                        d += ld_coh(G.dacc + ((size_t)buf * PT_DACC_SPLIT + sp) * Rp + Wl.rep);
                    E += d;
                }
                Wl.Eb = E;
                Wl.beta = in ? betaL[r] : 0.0;
                Wl.lab = in ? labelL[Wl.rep] : 0;
                Wl.trips = in ? (int32_t)stats[2 * Rp + Wl.rep] : 0;
                Wl.tried = in ? (int32_t)stats[r] : 0;
                Wl.acc = in ? (int32_t)stats[Rp + r] : 0;
                Wl.up = in ? (int32_t)stats[3 * Rp + r] : 0;
                Wl.down = in ? (int32_t)stats[4 * Rp + r] : 0;
                Wl.bestE = sh->best[0];
                Wl.best_round = (int32_t)sh->best[1];
                Wl.best_rung = (int32_t)sh->best[2];
                Wl.best_slot = (int32_t)sh->best[3];
                Wl.have = (int32_t)sh->best[4];
                // round rnd's end is round end rnd + 1 of the ladder (0 = the initial state)
                const bool imp = wave_ladder_round(Wl, I, rnd + 1, E, r, owner);
                if (in) {
                    repL[r] = Wl.rep;
                    rungL[Wl.rep] = r;
                    EbL[Wl.rep] = Wl.Eb;
                    labelL[Wl.rep] = Wl.lab;
                    stats[2 * Rp + Wl.rep] = Wl.trips;
                    stats[r] = Wl.tried;
                    stats[Rp + r] = Wl.acc;
                    stats[3 * Rp + r] = Wl.up;
                    stats[4 * Rp + r] = Wl.down;
                }
                if (r == 0) {
                    sh->best[0] = Wl.bestE;
                    sh->best[1] = Wl.best_round;
                    sh->best[2] = Wl.best_rung;
                    sh->best[3] = Wl.best_slot;
                    sh->best[4] = Wl.have;
                    sh->imp = imp;
                    sh->slot = Wl.best_slot;
                }
            }
            if (owner) {
                // buffer of round rnd + 2: its last readers finished before this round's barriers
                int64_t* z = G.dacc + (size_t)((rnd + 2) % 3) * PT_DACC_SPLIT * Rp;
                for (int x = tid; x < PT_DACC_SPLIT * Rp; x += nt)
                    __hip_atomic_store(z + x, (int64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            if (sh->imp) {
                // this workgroup's positions of the best slot, from its own LDS copy
                const int slot = sh->slot;
                int ob = 0;
                for (int c = 0; c < k; ++c) {
                    int lo, cnt;
                    slice(cptr[c], cptr[c + 1] - cptr[c], w, W, lo, cnt);
                    for (int i = tid; i < cnt; i += nt) I.best_s[lo + i] = lsp[(size_t)(ob + i) * Rp + slot];
                    ob += cnt;
                }
                __syncthreads();
            }
            continue;
        }
        for (int x = tid; x < R; x += nt) Es[x] = EbL[x];
        __syncthreads();
        // one coherent load per (counter, slot) pair, summed in LDS: a single memory round trip
        for (int idx = tid; idx < PT_DACC_SPLIT * R; idx += nt) {
            const int sp = idx / R, x = idx - sp * R;
            const int64_t v = ld_coh(G.dacc + ((size_t)buf * PT_DACC_SPLIT + sp) * Rp + x);
            if (v != 0) atomicAdd((unsigned long long*)&Es[x], (unsigned long long)v);
        }
        __syncthreads();
        int64_t mE;
        int32_t mR;
        block_best_rung(Es, repL, R, scrE, scrR, mE, mR);
        const bool improved = sh->best[4] == 0 || mE < sh->best[0];
        const int slot = repL[mR];
        if (improved) {
            // this workgroup's positions of the best slot, from its own LDS copy
            int ob = 0;
            for (int c = 0; c < k; ++c) {
                int lo, cnt;
                slice(cptr[c], cptr[c + 1] - cptr[c], w, W, lo, cnt);
                for (int i = tid; i < cnt; i += nt) I.best_s[lo + i] = lsp[(size_t)(ob + i) * Rp + slot];
                ob += cnt;
            }
        }
        if (owner) {
            if (tid == 0 && I.hist) I.hist[rnd] = (double)(improved ? mE : sh->best[0]) * I.unit;
            if (I.trace)
                for (int r = tid; r < R; r += nt)
                    I.trace[(size_t)rnd * R + r] = (double)Es[repL[r]] * I.unit;
        }
        __syncthreads();   // best and rep are read above before they change below
        if (tid == 0 && improved) {
            sh->best[0] = mE;
            sh->best[1] = rnd + 1;
            sh->best[2] = mR;
            sh->best[3] = slot;
            sh->best[4] = 1;
        }
        exchange_pairs(I, betaL, rnd, Es, repL, rungL, owner ? stats : nullptr,
                       owner ? stats + Rp : nullptr);
        __syncthreads();
        update_labels(R, rungL, labelL, nullptr, owner ? stats + 2 * Rp : nullptr,
                      owner ? stats + 3 * Rp : nullptr, owner ? stats + 4 * Rp : nullptr);
        for (int x = tid; x < R; x += nt) EbL[x] = Es[x];
        if (owner) {
            // buffer of round rnd + 2: its last readers finished before this round's barriers
            int64_t* z = G.dacc + (size_t)((rnd + 2) % 3) * PT_DACC_SPLIT * Rp;
            for (int x = tid; x < PT_DACC_SPLIT * Rp; x += nt)
                __hip_atomic_store(z + x, (int64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }

    // final state in the buffers of parity n_rounds & 1, where the host reads it
    if (owner) {
        const int fin = I.n_rounds & 1;
        for (int x = tid; x < R; x += nt) {
            G.Eb[(size_t)fin * Rp + x] = EbL[x];
            G.rung[(size_t)fin * Rp + x] = rungL[x];
            G.rep[(size_t)fin * Rp + x] = repL[x];
            G.label[x] = labelL[x];
            pt_out_arr(I.out, R, PT_OUT_TRIED)[x] = stats[x];
            pt_out_arr(I.out, R, PT_OUT_ACC)[x] = stats[Rp + x];
            pt_out_arr(I.out, R, PT_OUT_TRIPS)[x] = stats[2 * Rp + x];
            pt_out_arr(I.out, R, PT_OUT_UP)[x] = stats[3 * Rp + x];
            pt_out_arr(I.out, R, PT_OUT_DOWN)[x] = stats[4 * Rp + x];
        }
        if (tid < 5) G.best[5 * fin + tid] = sh->best[tid];
    }
}

template <int WM, bool UNI>
const void* coop_fn() {
    return reinterpret_cast<const void*>(&coop_run<WM, UNI>);
}

const void* coop_pick(int wm, bool uni) {
    if (wm == PT_WM_I8) return uni ? coop_fn<PT_WM_I8, true>() : coop_fn<PT_WM_I8, false>();
    if (wm == PT_WM_I32) return uni ? coop_fn<PT_WM_I32, true>() : coop_fn<PT_WM_I32, false>();
    return uni ? coop_fn<PT_WM_F32, true>() : coop_fn<PT_WM_F32, false>();
}

}  // namespace

hipError_t pt_coop_occupancy(int wm, bool uni, int block, size_t lds, int* per_cu) {
    const void* f = coop_pick(wm, uni);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, block, lds);
}

hipError_t pt_launch_coop(const PtInstance& I, const GState& G, const CoopArgs& C, int block,
                          int grid, size_t lds, bool cooperative, hipStream_t s) {
    PtInstance i = I;
    GState g = G;
    CoopArgs c = C;
    void* args[] = {(void*)&i, (void*)&g, (void*)&c};
    const void* f = coop_pick(I.wm, I.Rp % 64 == 0);
    return cooperative ? hipLaunchCooperativeKernel(f, dim3(grid), dim3(block), args, (unsigned)lds, s)
                       : hipLaunchKernel(f, dim3(grid), dim3(block), args, lds, s);
}
