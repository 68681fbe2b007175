// pt_slab.hip -- slab engine: one kernel per sweep for periodic lattices with +/-1 bonds.
//
// The lattice is cut into NB slabs of planes along its slowest axis (z in 3-D, y in 2-D) and one
// workgroup owns one slab of one replica.  It stages its planes plus two halo planes on each side
// in LDS, updates colour 0 on its planes and on one halo plane per side, then colour 1 on its own
// planes, and writes its planes to the other spin buffer.  The halo updates repeat exactly what the
// neighbouring workgroup computes for those sites -- the random stream is addressed by
// (position, replica, sweep), not by who computes it -- so no workgroup ever waits for another
// inside a launch, and the sweep needs one launch instead of one per colour.
//
// The first sweep kernel of round r also finishes round r-1: every workgroup recomputes the
// energies, the best-rung scan and the exchange decisions from a few hundred bytes and takes its
// own replica's new rung; workgroup 0 alone writes the bookkeeping.  Workgroups of one replica are
// placed on one XCD (blockIdx mod 8) so their halo reads hit the same L2.
//
// Positions follow the parity colouring in index order: position p of colour c is the site 2r or
// 2r+1 (r = p - c N/2) whose coordinate parity is c.  Reference counterpart: the sweep and swap
// kernels of cuda/ising_pt.cu:86-146.
#include "pt_device.h"
#include "pt_engine.h"
#include "pt_rules.h"

#include <algorithm>

namespace {

using namespace pt_dev;

struct Geo {
    int L, P, N, half, hp;   // side, sites per plane, sites, N/2, P/2
};

template <int DIM>
__device__ __forceinline__ Geo make_geo(const PtInstance& I) {
    Geo g;
    g.L = I.lL;
    g.N = I.n;
    g.P = DIM == 3 ? g.L * g.L : g.L;
    g.half = g.N >> 1;
    g.hp = g.P >> 1;
    return g;
}

// site holding position p of colour c
template <int DIM>
__device__ __forceinline__ int site_of(int p, int c, const Geo& g) {
    const int s0 = 2 * (p - c * g.half);
    const int yz = DIM == 3 ? (s0 / g.L) % g.L + s0 / g.P : s0 / g.L;
    return s0 + ((yz + c) & 1);
}

// colour-list position of a site (the slab engines need a power-of-two side, so coordinates
// come from shifts)
template <int DIM>
__device__ __forceinline__ int pos_of(int site, const Geo& g) {
    const int Lb = __builtin_ctz(g.L);
    const int par = DIM == 3 ? ((site & (g.L - 1)) + ((site >> Lb) & (g.L - 1)) + (site >> (2 * Lb))) & 1
                             : ((site & (g.L - 1)) + (site >> Lb)) & 1;
    return par * g.half + (site >> 1);
}

// workgroup -> (replica slot, slab); the NB slabs of one slot share blockIdx mod 8
__device__ __forceinline__ void wg_slot_slab(int w, int NB, int& a, int& b) {
    const int k = w >> 3;
    a = (k / NB) * 8 + (w & 7);
    b = k % NB;
}

// Round-end inputs of slot x = threadIdx.x, loaded before the sweep's staging so both sets of
// loads are in flight together.
struct ReIn {
    int64_t E;
    int32_t rep, rung;
    int64_t best[5];   // best record after the previous round end (tid 0 only)
};

__device__ __forceinline__ ReIn re_load(const PtInstance& I, const SState& S, int rnd, int x) {
    ReIn v{0, x, x, {0, 0, 0, 0, 0}};
    if (x == 0 && rnd > 0)
        for (int k = 0; k < 5; ++k) v.best[k] = S.best[5 * ((rnd - 1) & 1) + k];
    if (x >= I.R) return v;
    if (rnd == 0) {
        v.E = -(S.einit[x] / 2);
    } else {
        const int prev = rnd - 1, pin = prev & 1, din = (prev % 3 + 3) % 3;
        v.E = S.Eb[(size_t)pin * I.Rp + x] + S.dacc[(size_t)din * I.Rp + x];
        v.rep = S.rep[(size_t)pin * I.Rp + x];
        v.rung = S.rung[(size_t)pin * I.Rp + x];
    }
    return v;
}

// Round end of round rnd-1 (or the initial state when rnd == 0), evaluated by every workgroup.
// Returns the rung of slot a for round rnd.  mine holds re_load(tid); the best state of the
// workgroup's own planes is read as src[site - src_off] (the staged planes in LDS, or the spin
// buffer itself).
template <int DIM>
__device__ int slab_round_end(const PtInstance& I, const SState& S, const Geo& g, int rnd, int a,
                              const ReIn& mine_in, const int8_t* src, int src_off, int z0, int z1,
                              int64_t* Es, int32_t* reps, int32_t* rungs, long long* scrE,
                              int* scrR) {
    const int R = I.R, Rp = I.Rp, tid = threadIdx.x, nt = blockDim.x;
    const bool init = rnd == 0, owner = blockIdx.x == 0;
    const int prev = rnd - 1;
    if (tid < R) {
        Es[tid] = mine_in.E;
        reps[tid] = mine_in.rep;
        rungs[tid] = mine_in.rung;
    }
    for (int x = tid + nt; x < R; x += nt) {
        const ReIn v = re_load(I, S, rnd, x);
        Es[x] = v.E;
        reps[x] = v.rep;
        rungs[x] = v.rung;
    }
    __syncthreads();
    int64_t mE;
    int32_t mR;
    block_best_rung(Es, reps, R, scrE, scrR, mE, mR);
    __shared__ long long bcs[5];
    if (tid == 0)
        for (int k = 0; k < 5; ++k) bcs[k] = mine_in.best[k];
    __syncthreads();
    const long long* bc = bcs;
    const bool have = !init && bc[4] != 0;
    const bool improved = !have || mE < bc[0];
    const int slot = reps[mR];
    if (improved && slot == a)
        for (int site = z0 * g.P + tid; site < z1 * g.P; site += nt)
            I.best_s[pos_of<DIM>(site, g)] = src[site - src_off];
    int64_t* out = I.out;
    if (owner) {
        if (tid == 0) {
            int64_t* bn = S.best + 5 * (rnd & 1);
            if (improved) {
                bn[0] = mE; bn[1] = rnd; bn[2] = mR; bn[3] = slot; bn[4] = 1;
            } else {
                for (int x = 0; x < 5; ++x) bn[x] = bc[x];
            }
            if (!init && I.hist) I.hist[prev] = (double)(improved ? mE : bc[0]) * I.unit;
        }
        if (!init && I.trace)
            for (int r = tid; r < R; r += nt) I.trace[(size_t)prev * R + r] = (double)Es[reps[r]] * I.unit;
    }
    if (!init) {
        __syncthreads();
        exchange_pairs(I, prev, Es, reps, rungs, owner ? pt_out_arr(out, R, PT_OUT_TRIED) : nullptr,
                       owner ? pt_out_arr(out, R, PT_OUT_ACC) : nullptr);
        __syncthreads();
    }
    if (owner) {
        if (init) {
            for (int x = tid; x < R; x += nt)
                S.label[x] = (R > 1 && x == 0) ? 1 : ((R > 1 && x == R - 1) ? -1 : 0);
        } else {
            update_labels(R, rungs, S.label, nullptr, pt_out_arr(out, R, PT_OUT_TRIPS),
                          pt_out_arr(out, R, PT_OUT_UP), pt_out_arr(out, R, PT_OUT_DOWN));
        }
        const int cur = rnd & 1, zero = (rnd + 1) % 3;
        for (int x = tid; x < R; x += nt) {
            S.Eb[(size_t)cur * Rp + x] = Es[x];
            S.rung[(size_t)cur * Rp + x] = rungs[x];
            S.rep[(size_t)cur * Rp + x] = reps[x];
            S.dacc[(size_t)zero * Rp + x] = 0;
        }
    }
    const int mine = rungs[a];
    __syncthreads();
    return mine;
}

// Both colour steps of sweep t on a staged slab: colour 0 on the owned planes and one halo plane
// per side, then colour 1 on the owned planes.  Returns the energy change of the owned sites.
// lsp: spin slots (slot s holds plane (WRAP ? s : z0 - 2 + s) mod L); ljb: bond signs of slots
// jlo.. ; t2, t4, t6: thresholds of q = 2, 4, 6 at this replica's rung.
template <int DIM, bool WRAP>
__device__ __forceinline__ int64_t slab_colours(const PtInstance& I, const Geo& g, int8_t* lsp,
                                                const int8_t* ljb, int a, uint32_t t, int z0,
                                                int T, int jlo, uint32_t t2, uint32_t t4,
                                                uint32_t t6) {
    const int L = g.L, tid = threadIdx.x, nt = blockDim.x;
    const int Lb = 31 - __builtin_clz(L);
    const int Pb = DIM == 3 ? 2 * Lb : Lb;
    int64_t dsum = 0;
    // Quad k of plane `plane`, colour c holds positions c N/2 + plane P/2 + 4k .. +3, i.e. the
    // colour-c site of each of the site pairs (plane P + 8k + 2kk, +1), kk = 0..3.
    const int Qb = Pb - 3;                   // log2 of quads per plane and colour (P/8)
    for (int c = 0; c < 2; ++c) {
        const int s_lo = WRAP ? 0 : (c == 0 ? 1 : 2);
        const int ns = WRAP ? L : (c == 0 ? T + 2 : T);
        for (int w = tid; w < (ns << Qb); w += nt) {
            const int s = s_lo + (w >> Qb), k = w & ((1 << Qb) - 1);
            const int plane = WRAP ? s : (z0 - 2 + s) & (L - 1);
            const bool owned = WRAP || (s >= 2 && s < T + 2);
            const int su = WRAP ? ((s + 1) & (L - 1)) : s + 1;
            const int sd = WRAP ? ((s - 1) & (L - 1)) : s - 1;
            const uint32_t q = (uint32_t)((c * g.half + (plane << (Pb - 1))) >> 2) + (uint32_t)k;
            const pt_u4 rv = pt_philox(q, (uint32_t)a, t, PT_TAG_METRO, I.k0, I.k1);
            const int8_t* ls = lsp + ((size_t)s << Pb);
            const int8_t* lu = lsp + ((size_t)su << Pb);
            const int8_t* ld = lsp + ((size_t)sd << Pb);
            const int8_t* lj = ljb + ((size_t)(s - jlo) << Pb);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int loc0 = 8 * k + 2 * kk;              // even site of the pair, in plane
                const int y = DIM == 3 ? loc0 >> Lb : plane;
                const int par = DIM == 3 ? ((y + plane + c) & 1) : ((plane + c) & 1);
                const int loc = loc0 + par;
                const int x = loc & (L - 1);
                const int rowb = loc - x;
                const uint32_t jb = (uint8_t)lj[loc];
                const int n0 = ls[rowb + ((x + 1) & (L - 1))];
                const int n1 = ls[rowb + ((x - 1) & (L - 1))];
                int l = ((jb & 1u) ? -n0 : n0) + ((jb & 2u) ? -n1 : n1);
                if (DIM == 3) {
                    const int n2 = ls[(((y + 1) & (L - 1)) << Lb) + x];
                    const int n3 = ls[(((y - 1) & (L - 1)) << Lb) + x];
                    const int n4 = lu[loc];
                    const int n5 = ld[loc];
                    l += ((jb & 4u) ? -n2 : n2) + ((jb & 8u) ? -n3 : n3) +
                         ((jb & 16u) ? -n4 : n4) + ((jb & 32u) ? -n5 : n5);
                } else {
                    const int n2 = lu[loc];
                    const int n3 = ld[loc];
                    l += ((jb & 4u) ? -n2 : n2) + ((jb & 8u) ? -n3 : n3);
                }
                int8_t* me = (int8_t*)ls + loc;
                const int8_t si = *me;
                const int qv = si > 0 ? l : -l;
                const uint32_t th = qv == 2 ? t2 : (qv == 4 ? t4 : t6);
                if (qv <= 0 || pt_word(rv, kk) < th) {
                    *me = (int8_t)-si;
                    if (owned) dsum += 2 * qv;
                }
            }
        }
        __syncthreads();
    }

    return dsum;
}

template <int DIM, bool WRAP>
__global__ __launch_bounds__(1024) void slab_sweep(PtInstance I, SState S, int j, int rl,
                                                   int do_sweep) {
    extern __shared__ __attribute__((aligned(16))) int8_t lsp[];
    __shared__ int64_t Es[1024];
    __shared__ int32_t reps[1024];
    __shared__ int32_t rungs[1024];
    __shared__ uint32_t thr3[3 * 1024];      // thresholds of q = 2, 4, 6 for every rung
    __shared__ long long scrE[16];
    __shared__ int scrR[16];
    int a, b;
    wg_slot_slab(blockIdx.x, S.NB, a, b);
    if (a >= I.R) return;   // whole workgroup: padding slots of the grid
    const Geo g = make_geo<DIM>(I);
    const int L = g.L, P = g.P, tid = threadIdx.x, nt = blockDim.x;
    const int Lb = 31 - __builtin_clz(L);                // L is a power of two
    const int Pb = DIM == 3 ? 2 * Lb : Lb;
    const int z0 = b * L / S.NB, z1 = (b + 1) * L / S.NB, T = z1 - z0;
    const int rnd = *S.round_base + rl;
    const uint32_t t = (uint32_t)rnd * (uint32_t)I.S + (uint32_t)j;
    const int8_t* in = S.sp + ((size_t)(t & 1) * I.Rp + a) * g.N;
    int8_t* outb = S.sp + ((size_t)((t + 1) & 1) * I.Rp + a) * g.N;
    const bool re = j == 0;
    const ReIn mine_in = re ? re_load(I, S, rnd, tid) : ReIn{0, 0, 0, {0, 0, 0, 0, 0}};
    if (!do_sweep) {
        if (re) slab_round_end<DIM>(I, S, g, rnd, a, mine_in, in, 0, z0, z1, Es, reps, rungs, scrE, scrR);
        return;
    }
    const int keep_rung = re ? 0 : S.rung[(size_t)(rnd & 1) * I.Rp + a];

    // Stage spins of slots 0..nslots-1 and bond signs of the updated slots 1..T+2 (all slots when
    // the lattice wraps inside the workgroup); slot s holds plane (WRAP ? s : z0 - 2 + s) mod L.
    // Each thread issues eight independent word loads per pass (an index past the end reads word
    // 0 and is not stored), then the thresholds of every rung.
    const int nslots = WRAP ? L : T + 4;
    const int jlo = WRAP ? 0 : 1, njs = WRAP ? L : T + 2;
    int8_t* ljb = lsp + (size_t)nslots * P;
    const int wpp = P >> 2;
    const int nsw = nslots * wpp, total = nsw + njs * wpp;
    for (int base = 0; base < total; base += 8 * nt) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int idx = base + u * nt + tid;
            idx = idx < total ? idx : 0;
            const bool spin = idx < nsw;
            const int r = spin ? idx : idx - nsw;
            const int s = (spin ? 0 : jlo) + (r >> (Pb - 2)), w = r & (wpp - 1);
            const int plane = WRAP ? s : (z0 - 2 + s) & (L - 1);
            const uint32_t* src = (const uint32_t*)((spin ? in : (const int8_t*)I.jbits) + ((size_t)plane << Pb));
            v[u] = src[w];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = base + u * nt + tid;
            if (idx < total) ((uint32_t*)lsp)[idx < nsw ? idx : idx - nsw + ((nslots * P) >> 2)] = v[u];
        }
    }
    for (int x = tid; x < I.R; x += nt) {
        const uint32_t* tr = I.thr + (size_t)x * I.TQ;
        thr3[3 * x] = I.TQ >= 2 ? tr[1] : 0u;
        thr3[3 * x + 1] = I.TQ >= 4 ? tr[3] : 0u;
        thr3[3 * x + 2] = (DIM == 3 && I.TQ >= 6) ? tr[5] : 0u;
    }
    __syncthreads();
    const int so = WRAP ? 0 : 2;
    const int g_rung = re ? slab_round_end<DIM>(I, S, g, rnd, a, mine_in, lsp + (size_t)so * P,
                                                z0 * P, z0, z1, Es, reps, rungs, scrE, scrR)
                          : keep_rung;
    const uint32_t t2 = thr3[3 * g_rung], t4 = thr3[3 * g_rung + 1], t6 = thr3[3 * g_rung + 2];

    const int64_t dsum = slab_colours<DIM, WRAP>(I, g, lsp, ljb, a, t, z0, T, jlo, t2, t4, t6);

    // own planes to the other buffer
    for (int idx = tid; idx < T * wpp; idx += nt) {
        const int s = idx >> (Pb - 2), w = idx & (wpp - 1);
        ((uint32_t*)(outb + ((size_t)(z0 + s) << Pb)))[w] = ((const uint32_t*)lsp)[((s + so) << (Pb - 2)) + w];
    }
    // energy change of the owned sites: one atomic per workgroup
    long long v = (long long)dsum;
#pragma unroll
    for (int m = PT_WAVE / 2; m > 0; m >>= 1) v += __shfl_xor(v, m);
    if ((tid & (PT_WAVE - 1)) == 0) scrE[tid / PT_WAVE] = v;
    __syncthreads();
    if (tid == 0) {
        long long tot = 0;
        for (int w = 0; w < (nt + PT_WAVE - 1) / PT_WAVE; ++w) tot += scrE[w];
        if (tot != 0)
            atomicAdd((unsigned long long*)&S.dacc[(size_t)(rnd % 3) * I.Rp + a], (unsigned long long)tot);
    }
}

// initial spins of every slot (buffer 0, site order) from the stream contract
__global__ __launch_bounds__(256) void slab_init_spins(PtInstance I, SState S) {
    const int64_t total = (int64_t)I.R * I.n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(idx / I.n), site = (int)(idx - (int64_t)a * I.n);
        const pt_u4 r = pt_philox((uint32_t)(site >> 2), (uint32_t)a, 0u, PT_TAG_INIT, I.k0, I.k1);
        S.sp[(size_t)a * I.n + site] = (pt_word(r, site & 3) & 1u) ? (int8_t)1 : (int8_t)-1;
    }
}

// einit[a] = sum_i s_i sum_j J_ij s_j over all sites (E = -einit/2); one slot per block row
template <int DIM>
__global__ __launch_bounds__(256) void slab_init_energy(PtInstance I, SState S) {
    __shared__ long long scr[8];
    const int a = blockIdx.y;
    const int L = I.lL, P = DIM == 3 ? L * L : L;
    const int8_t* sp = S.sp + (size_t)a * I.n;
    long long acc = 0;
    for (int site = blockIdx.x * blockDim.x + threadIdx.x; site < I.n; site += gridDim.x * blockDim.x) {
        const int x = site % L, y = (site / L) % L, z = DIM == 3 ? site / P : 0;
        const int base = DIM == 3 ? z * P : 0;
        const uint32_t jb = I.jbits[site];
        int nb[6];
        nb[0] = base + y * L + (x + 1) % L;
        nb[1] = base + y * L + (x + L - 1) % L;
        nb[2] = base + ((y + 1) % L) * L + x;
        nb[3] = base + ((y + L - 1) % L) * L + x;
        if (DIM == 3) {
            nb[4] = ((z + 1) % L) * P + y * L + x;
            nb[5] = ((z + L - 1) % L) * P + y * L + x;
        }
        int l = 0;
        for (int d = 0; d < 2 * DIM; ++d) {
            const int v = sp[nb[d]];
            l += ((jb >> d) & 1u) ? -v : v;
        }
        acc += sp[site] > 0 ? l : -l;
    }
#pragma unroll
    for (int m = PT_WAVE / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m);
    if ((threadIdx.x & (PT_WAVE - 1)) == 0) scr[threadIdx.x / PT_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tot = 0;
        for (int w = 0; w < (int)(blockDim.x / PT_WAVE); ++w) tot += scr[w];
        atomicAdd((unsigned long long*)&S.einit[a], (unsigned long long)tot);
    }
}

__global__ void slab_tick(int32_t* round_base, int rounds) {
    if (threadIdx.x == 0) *round_base += rounds;
}

// Persistent slab engine: one cooperative launch runs every round.  Each workgroup keeps its
// slab (owned planes, two halo planes per side, bond signs) in LDS for the whole run.  After a
// sweep it publishes its two first and two last owned planes into the spin buffer of the next
// sweep parity (agent-scope stores); after the grid barrier it reads its four halo planes from
// there (agent-scope loads).  Round ends are evaluated identically by every workgroup from the
// energy sums; workgroup 0 alone writes the statistics.  Padding workgroups (slots >= R) only
// take part in the barriers.
template <int DIM, bool WRAP>
__global__ __launch_bounds__(1024) void pslab_run(PtInstance I, SState S, CoopArgs C) {
    extern __shared__ __attribute__((aligned(16))) int8_t lsp[];
    __shared__ int64_t Es[1024];
    __shared__ int64_t EbL[1024];
    __shared__ int32_t reps[1024];
    __shared__ int32_t rungs[1024];
    __shared__ uint32_t thr3[3 * 1024];
    __shared__ long long scrE[16];
    __shared__ int scrR[16];
    __shared__ long long bestL[5];
    __shared__ double betaL[1024];
    __shared__ int32_t labelL[1024];   // round-trip labels (workgroup 0's copy is the one used)
    __shared__ int sh_abort, sh_imp, sh_slot;
    __shared__ long long sh_dsum;
    int a, b;
    wg_slot_slab(blockIdx.x, S.NB, a, b);
    const bool live = a < I.R, owner = blockIdx.x == 0;
    const Geo g = make_geo<DIM>(I);
    const int L = g.L, P = g.P, tid = threadIdx.x, nt = blockDim.x;
    const int Lb = 31 - __builtin_clz(L);
    const int Pb = DIM == 3 ? 2 * Lb : Lb;
    const int R = I.R, Rp = I.Rp;
    const int z0 = b * L / S.NB, z1 = (b + 1) * L / S.NB, T = z1 - z0;
    const int nslots = WRAP ? L : T + 4;
    const int jlo = WRAP ? 0 : 1, njs = WRAP ? L : T + 2;
    const int so = WRAP ? 0 : 2;
    int8_t* ljb = lsp + (size_t)nslots * P;
    const int wpp = P >> 2;

    // ---- stage the slab from spin buffer 0 (written by the init kernels) ----------------------
    if (live) {
        const int8_t* in = S.sp + (size_t)a * g.N;
        for (int idx = tid; idx < nslots * wpp; idx += nt) {
            const int sl = idx >> (Pb - 2), w = idx & (wpp - 1);
            const int plane = WRAP ? sl : (z0 - 2 + sl) & (L - 1);
            ((uint32_t*)lsp)[idx] = ((const uint32_t*)(in + ((size_t)plane << Pb)))[w];
        }
        for (int idx = tid; idx < njs * wpp; idx += nt) {
            const int sl = jlo + (idx >> (Pb - 2)), w = idx & (wpp - 1);
            const int plane = WRAP ? sl : (z0 - 2 + sl) & (L - 1);
            ((uint32_t*)ljb)[idx] = ((const uint32_t*)(I.jbits + ((size_t)plane << Pb)))[w];
        }
    }
    for (int x = tid; x < R; x += nt) {
        betaL[x] = I.beta[x];
        labelL[x] = (R > 1 && x == 0) ? 1 : ((R > 1 && x == R - 1) ? -1 : 0);
        const uint32_t* tr = I.thr + (size_t)x * I.TQ;
        thr3[3 * x] = I.TQ >= 2 ? tr[1] : 0u;
        thr3[3 * x + 1] = I.TQ >= 4 ? tr[3] : 0u;
        thr3[3 * x + 2] = (DIM == 3 && I.TQ >= 6) ? tr[5] : 0u;
    }
    if (tid == 0) {
        sh_abort = 0;
        sh_dsum = 0;
        for (int k = 0; k < 5; ++k) bestL[k] = 0;
    }
    // R <= 64: the round end is wave 0's (WaveLadder); the other waves meanwhile bring in the
    // halo planes of the round's last sweep
    const bool fast = R <= PT_WAVE;
    const bool defer = fast && !WRAP && nt > PT_WAVE && I.S > 0;
    WaveLadder WL;
    if (fast && tid < PT_WAVE) wave_ladder_init(WL, I, S.einit, tid);
    __syncthreads();

    uint32_t phase = 0;
    for (int rnd = 0;; ++rnd) {
        // ---- end of round rnd - 1 (rnd == 0: the initial state), identical in every workgroup
        if (fast) {
            if (tid < PT_WAVE) {
                int64_t E = WL.Eb;
                if (rnd > 0 && tid < R)
                    E += ld_coh(S.dacc + (size_t)((rnd - 1) % 3) * Rp + WL.rep);
                const bool imp = wave_ladder_round(WL, I, rnd, E, tid, owner);
                if (tid < R) rungs[WL.rep] = tid;
                if (tid == 0) {
                    sh_imp = imp;
                    sh_slot = WL.best_slot;
                }
                if (owner && tid < R)   // buffer of round rnd + 1 (see below)
                    __hip_atomic_store(S.dacc + (size_t)((rnd + 1) % 3) * Rp + tid, (int64_t)0,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (defer && live && rnd > 0 && rnd < I.n_rounds) {
                const int8_t* prv = S.sp + ((size_t)(((uint32_t)rnd * (uint32_t)I.S) & 1) * Rp + a) * g.N;
                for (int idx = tid - PT_WAVE; idx < 4 * wpp; idx += nt - PT_WAVE) {
                    const int k = idx >> (Pb - 2), w = idx & (wpp - 1);
                    const int sl = k < 2 ? k : T + k;
                    const int plane = (z0 - 2 + sl) & (L - 1);
                    ((uint32_t*)lsp)[(sl << (Pb - 2)) + w] =
                        ld_u32((const uint32_t*)(prv + ((size_t)plane << Pb)) + w);
                }
            }
            __syncthreads();
            if (sh_imp) {
                if (live && sh_slot == a)
                    for (int site = z0 * P + tid; site < z1 * P; site += nt)
                        st_spin<true>(I.best_s + pos_of<DIM>(site, g), lsp[(size_t)so * P + site - z0 * P]);
                __syncthreads();   // the spins are read above before the sweep changes them
            }
            if (rnd == I.n_rounds) break;
        } else {
        for (int x = tid; x < R; x += nt) {
            if (rnd == 0) {
                Es[x] = -(S.einit[x] / 2);
                reps[x] = x;
                rungs[x] = x;
            } else {
                Es[x] = EbL[x] + ld_coh(S.dacc + (size_t)((rnd - 1) % 3) * Rp + x);
            }
        }
        __syncthreads();
        int64_t mE;
        int32_t mR;
        block_best_rung(Es, reps, R, scrE, scrR, mE, mR);
        const bool improved = bestL[4] == 0 || mE < bestL[0];
        const int slot = reps[mR];
        // best state of the owned planes.  Different rounds' best slots live on different XCDs,
        // so these bytes are written through (agent scope): plain stores would leave dirty copies
        // in several L2s, and the last write-back at kernel end -- not the last round -- would win.
        if (improved && live && slot == a)
            for (int site = z0 * P + tid; site < z1 * P; site += nt)
                st_spin<true>(I.best_s + pos_of<DIM>(site, g), lsp[(size_t)so * P + site - z0 * P]);
        if (owner && rnd > 0) {
            if (tid == 0 && I.hist) I.hist[rnd - 1] = (double)(improved ? mE : bestL[0]) * I.unit;
            if (I.trace)
                for (int r = tid; r < R; r += nt)
                    I.trace[(size_t)(rnd - 1) * R + r] = (double)Es[reps[r]] * I.unit;
        }
        __syncthreads();   // best and reps are read above before they change below
        if (tid == 0 && improved) {
            bestL[0] = mE;
            bestL[1] = rnd;
            bestL[2] = mR;
            bestL[3] = slot;
            bestL[4] = 1;
        }
        if (rnd > 0) {
            exchange_pairs(I, betaL, rnd - 1, Es, reps, rungs,
                           owner ? pt_out_arr(I.out, R, PT_OUT_TRIED) : nullptr,
                           owner ? pt_out_arr(I.out, R, PT_OUT_ACC) : nullptr);
            __syncthreads();
            if (owner)
                update_labels(R, rungs, labelL, nullptr, pt_out_arr(I.out, R, PT_OUT_TRIPS),
                              pt_out_arr(I.out, R, PT_OUT_UP), pt_out_arr(I.out, R, PT_OUT_DOWN));
        }
        for (int x = tid; x < R; x += nt) EbL[x] = Es[x];
        if (owner)   // buffer of round rnd + 1: its last readers finished before this round began
            for (int x = tid; x < R; x += nt)
                __hip_atomic_store(S.dacc + (size_t)((rnd + 1) % 3) * Rp + x, (int64_t)0,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (rnd == I.n_rounds) break;
        }

        // ---- the sweeps of round rnd --------------------------------------------------------
        const int gr = live ? rungs[a] : 0;
        const uint32_t t2 = thr3[3 * gr], t4 = thr3[3 * gr + 1], t6 = thr3[3 * gr + 2];
        int64_t dsum = 0;
        for (int j = 0; j < I.S; ++j) {
            const uint32_t t = (uint32_t)rnd * (uint32_t)I.S + (uint32_t)j;
            if (live) dsum += slab_colours<DIM, WRAP>(I, g, lsp, ljb, a, t, z0, T, jlo, t2, t4, t6);
            int8_t* nxt = S.sp + ((size_t)((t + 1) & 1) * Rp + a) * g.N;
            if (live && !WRAP) {
                // the first and last min(2, T) owned planes: the neighbours' halos
                const int nb = T < 2 ? T : 2;
                for (int idx = tid; idx < 2 * nb * wpp; idx += nt) {
                    const int k = idx >> (Pb - 2), w = idx & (wpp - 1);
                    const int o = k < nb ? k : T - 2 * nb + k;
                    st_u32((uint32_t*)(nxt + ((size_t)(z0 + o) << Pb)) + w,
                           ((const uint32_t*)lsp)[((so + o) << (Pb - 2)) + w]);
                }
            }
            const bool last = j == I.S - 1;
            if (last) {   // the slot's energy change: waves into LDS, thread 0 to HBM on arrival
                long long v = (long long)dsum;
#pragma unroll
                for (int m = PT_WAVE / 2; m > 0; m >>= 1) v += __shfl_xor(v, m);
                if ((tid & (PT_WAVE - 1)) == 0 && v != 0 && live)
                    atomicAdd((unsigned long long*)&sh_dsum, (unsigned long long)v);
            }
            if (!coop_sync(C, phase, &sh_abort, S.dacc + (size_t)(rnd % 3) * Rp + a,
                           last && live ? &sh_dsum : nullptr))
                return;
            if (last && defer) continue;   // halo planes come in during the round end
            if (live && !WRAP) {
                // halo slots 0, 1, T + 2, T + 3 as the neighbours left them after this sweep
                for (int idx = tid; idx < 4 * wpp; idx += nt) {
                    const int k = idx >> (Pb - 2), w = idx & (wpp - 1);
                    const int sl = k < 2 ? k : T + k;
                    const int plane = (z0 - 2 + sl) & (L - 1);
                    ((uint32_t*)lsp)[(sl << (Pb - 2)) + w] =
                        ld_u32((const uint32_t*)(nxt + ((size_t)plane << Pb)) + w);
                }
            }
            __syncthreads();
        }
    }

    // ---- final state where the host reads it ---------------------------------------------------
    if (live) {
        int8_t* fb = S.sp + ((size_t)(((int64_t)I.n_rounds * I.S) & 1) * Rp + a) * g.N;
        for (int idx = tid; idx < T * wpp; idx += nt) {
            const int o = idx >> (Pb - 2), w = idx & (wpp - 1);
            ((uint32_t*)(fb + ((size_t)(z0 + o) << Pb)))[w] = ((const uint32_t*)lsp)[((so + o) << (Pb - 2)) + w];
        }
    }
    if (owner && fast) {
        const int fin = I.n_rounds & 1;
        if (tid < PT_WAVE)
            wave_ladder_store(WL, I, S.Eb + (size_t)fin * Rp, S.rung + (size_t)fin * Rp,
                              S.rep + (size_t)fin * Rp, S.label, S.best + 5 * fin, tid);
    } else if (owner) {
        const int fin = I.n_rounds & 1;
        for (int x = tid; x < R; x += nt) {
            S.Eb[(size_t)fin * Rp + x] = EbL[x];
            S.rung[(size_t)fin * Rp + x] = rungs[x];
            S.rep[(size_t)fin * Rp + x] = reps[x];
            S.label[x] = labelL[x];
        }
        if (tid < 5) S.best[5 * fin + tid] = bestL[tid];
    }
}

// ---------------------------------------------------------------------------------------------
// Bit-sliced persistent slab engine (side L a multiple of 32).  Same workgroup layout, rounds,
// barriers and round ends as pslab_run, but the slab lives in LDS as bits: row words of 32 sites
// (1 = spin -1) and one bit-plane per bond direction (1 = bond -1).  A lane takes an 8-site
// segment of a row: it XORs the six neighbour words (the row shifted by one site either way, the
// rows above and below, the same row in the planes above and below) with the bond planes,
// counts the unsatisfied terms of all 32 sites at once with full adders, and then decides the
// segment's four sites of the colour being updated against its Philox quad -- the uniforms every
// engine draws -- and flips them with one LDS atomic XOR.  Halo planes move between workgroups
// as bit planes.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void fadd(uint32_t a, uint32_t b, uint32_t c, uint32_t& s, uint32_t& cy) {
    const uint32_t t = a ^ b;
    s = t ^ c;
    cy = (a & b) | (c & t);
}

// int8 spins of one row word (32 sites) -> bits, 1 for spin -1
__device__ __forceinline__ uint32_t row_bits(const int8_t* src) {
    uint32_t bits = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint32_t v = ((const uint32_t*)src)[q];
        const uint32_t nib = ((v >> 7) & 1u) | ((v >> 14) & 2u) | ((v >> 21) & 4u) | ((v >> 28) & 8u);
        bits |= nib << (4 * q);
    }
    return bits;
}

// bit d of each of 32 bond-sign bytes -> one word
__device__ __forceinline__ uint32_t bond_bits(const uint8_t* src, int d) {
    uint32_t bits = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint32_t v = ((const uint32_t*)src)[q] >> d;
        const uint32_t nib = (v & 1u) | ((v >> 7) & 2u) | ((v >> 14) & 4u) | ((v >> 21) & 8u);
        bits |= nib << (4 * q);
    }
    return bits;
}

// bits -> 32 int8 spins
__device__ __forceinline__ void bits_row(uint32_t bits, int8_t* dst) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        uint32_t v = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) v |= (((bits >> (4 * q + e)) & 1u) ? 0xFFu : 0x01u) << (8 * e);
        ((uint32_t*)dst)[q] = v;
    }
}

// Geometry and LDS layout of one workgroup's bit slab.  Slot s of the slab holds plane
// (WRAP ? s : z0 - 2 + s) mod L.  A sweep is two passes: colour 0 over the slab and its two halo
// planes (slots 1 .. T + 2), then colour 1 over the owned planes (slots 2 .. T + 1).
struct BitSlab {
    int L, P, Lb, WPR, WPP, wppb;          // side, sites per plane, log2 L, words per row and plane,
                                           // log2 words per plane
    int z0, T, nslots, jlo, njs, so;       // first owned plane, owned planes, staged spin slots,
                                           // first bond slot, bond slots, slot of plane z0
    int segb, rb, wb, pb, psb, jstride;    // log2 of: 8-site segments per row, rows per plane, words
                                           // per row, sites per plane, segments per slot; words per
                                           // bond direction
    int ns0, ns1, uoff1, nU;               // slots of the two passes; uniforms of a sweep
    uint32_t* sb;                          // [slot][row][word] spin bits (1 = spin -1)
    uint32_t* jb;                          // [dir][bond slot][row][word] bond bits (1 = bond -1)
    uint4* U;                              // next-sweep uniforms: one Philox block per segment
};

template <int DIM, bool WRAP>
__device__ __forceinline__ BitSlab make_bit_slab(const Geo& g, int NB, int b, uint32_t* lw) {
    BitSlab B;
    B.L = g.L;
    B.P = g.P;
    B.Lb = 31 - __builtin_clz(g.L);
    B.WPR = g.L >> 5;
    B.WPP = (DIM == 3 ? g.L : 1) * B.WPR;
    B.wppb = (DIM == 3 ? B.Lb : 0) + B.Lb - 5;
    B.z0 = b * g.L / NB;
    B.T = (b + 1) * g.L / NB - B.z0;
    B.nslots = WRAP ? g.L : B.T + 4;
    B.jlo = WRAP ? 0 : 1;
    B.njs = WRAP ? g.L : B.T + 2;
    B.so = WRAP ? 0 : 2;
    B.segb = B.Lb - 3;
    B.rb = DIM == 3 ? B.Lb : 0;
    B.wb = B.Lb - 5;
    B.pb = DIM == 3 ? 2 * B.Lb : B.Lb;
    B.psb = B.rb + B.segb;
    B.jstride = B.njs << B.wppb;
    B.ns0 = WRAP ? g.L : B.T + 2;
    B.ns1 = WRAP ? g.L : B.T;
    B.uoff1 = B.ns0 << B.psb;
    B.nU = B.uoff1 + (B.ns1 << B.psb);
    B.sb = lw;
    B.jb = lw + (size_t)B.nslots * B.WPP;
    B.U = (uint4*)(lw + ((((size_t)B.nslots + (size_t)2 * DIM * B.njs) * B.WPP + 3) & ~(size_t)3));
    return B;
}

// Pass c, slot i of the pass -> slab slot.
template <bool WRAP>
__device__ __forceinline__ int bit_pass_slot(int c, int i) {
    return WRAP ? i : (c == 0 ? 1 + i : 2 + i);
}

// Stage slot a's slab (spin buffer 0, written by the init kernels) and its bond planes as bits.
template <int DIM, bool WRAP>
__device__ void bit_stage(const BitSlab& B, const PtInstance& I, const SState& S, const Geo& g,
                          int a) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int8_t* in = S.sp + (size_t)a * g.N;
    for (int idx = tid; idx < B.nslots * B.WPP; idx += nt) {
        const int sl = idx >> B.wppb, rw = idx & (B.WPP - 1);
        const int plane = WRAP ? sl : (B.z0 - 2 + sl) & (B.L - 1);
        B.sb[idx] = row_bits(in + ((size_t)plane * B.P) + ((size_t)rw << 5));
    }
    for (int idx = tid; idx < 2 * DIM * B.njs * B.WPP; idx += nt) {
        const int d = idx / (B.njs * B.WPP), r2 = idx - d * B.njs * B.WPP;
        const int sl = B.jlo + (r2 >> B.wppb), rw = r2 & (B.WPP - 1);
        const int plane = WRAP ? sl : (B.z0 - 2 + sl) & (B.L - 1);
        B.jb[idx] = bond_bits(I.jbits + ((size_t)plane * B.P) + ((size_t)rw << 5), d);
    }
}

// The Philox blocks of sweep tt (they depend on position, slot and sweep only) into U, entries
// first, first + stride, ...
template <int DIM, bool WRAP>
__device__ void bit_draw(const BitSlab& B, const PtInstance& I, const Geo& g, int a, uint32_t tt,
                         int first, int stride) {
    for (int idx = first; idx < B.nU; idx += stride) {
        const int c = idx < B.uoff1 ? 0 : 1;
        const int w = idx - (c == 0 ? 0 : B.uoff1);
        const int i = w >> B.psb, rem = w & ((1 << B.psb) - 1);
        const int s = bit_pass_slot<WRAP>(c, i);
        const int r = rem >> B.segb, k = rem & ((1 << B.segb) - 1);
        const int plane = WRAP ? s : (B.z0 - 2 + s) & (B.L - 1);
        const uint32_t q = ((uint32_t)(c * g.half) >> 2) +
                           ((uint32_t)((plane << B.pb) + (r << B.Lb) + (k << 3)) >> 3);
        const pt_u4 v = pt_philox(q, (uint32_t)a, tt, PT_TAG_METRO, I.k0, I.k1);
        B.U[idx] = make_uint4(v.x, v.y, v.z, v.w);
    }
}

// Both colour passes of sweep t (every thread of the workgroup calls it; each pass ends with a
// workgroup barrier).  Returns the energy change of the owned sites.
//
// One lane per 32-site row word: the neighbour words, bond planes and unsatisfied-bond counts of
// the word serve its 16 sites of the colour, decided against four Philox blocks (one per 8-site
// segment, the blocks every engine draws; from U when pre).  The lane owns the word in this pass;
// lanes reading it as a neighbour use only the other colour's bits, which do not change, so a
// plain store suffices.  t2, t4, t6: thresholds of q = 2, 4, 6 at the slot's rung.
template <int DIM, bool WRAP>
__device__ __forceinline__ int64_t bit_sweep(const BitSlab& B, const PtInstance& I, const Geo& g,
                                             int a, uint32_t t, bool pre, uint32_t t2, uint32_t t4,
                                             uint32_t t6) {
    constexpr int D = 2 * DIM;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int L = B.L, Lb = B.Lb, rb = B.rb, wb = B.wb, WPR = B.WPR, T = B.T;
    uint32_t* sb = B.sb;
    const uint32_t* jb = B.jb;
    const int wpb = rb + wb;                      // log2 words per slot
    int64_t dsum = 0;
    for (int c = 0; c < 2; ++c) {
        const int ns = c == 0 ? B.ns0 : B.ns1;
        const int uo = c == 0 ? 0 : B.uoff1;
        for (int w = tid; w < (ns << wpb); w += nt) {
            const int i = w >> wpb, rem = w & ((1 << wpb) - 1);
            const int s = bit_pass_slot<WRAP>(c, i);
            const int r = rem >> wb, wi = rem & (WPR - 1);
            const int plane = WRAP ? s : (B.z0 - 2 + s) & (L - 1);
            const bool owned = WRAP || (s >= 2 && s < T + 2);
            const int su = WRAP ? ((s + 1) & (L - 1)) : s + 1;
            const int sd = WRAP ? ((s - 1) & (L - 1)) : s - 1;
            // word index of (slot, row) + wi, by shifts: rows = 2^rb, words/row = 2^wb
            const int wrow = (((s << rb) + r) << wb) + wi;
            const uint32_t W = sb[wrow];
            const uint32_t Wn = sb[wrow - wi + ((wi + 1) & (WPR - 1))];
            const uint32_t Wp = sb[wrow - wi + ((wi - 1) & (WPR - 1))];
            uint32_t A[6];
            A[0] = (W >> 1) | (Wn << 31);               // +x
            A[1] = (W << 1) | (Wp >> 31);               // -x
            if (DIM == 3) {
                A[2] = sb[(((s << rb) + ((r + 1) & (L - 1))) << wb) + wi];
                A[3] = sb[(((s << rb) + ((r - 1) & (L - 1))) << wb) + wi];
                A[4] = sb[(((su << rb) + r) << wb) + wi];
                A[5] = sb[(((sd << rb) + r) << wb) + wi];
            } else {
                A[2] = sb[(su << wb) + wi];
                A[3] = sb[(sd << wb) + wi];
            }
            const int jrow = ((((s - B.jlo) << rb) + r) << wb) + wi;
            // Y[d] = 1 where the bond to neighbour d is unsatisfied (s_i J s_j = -1):
            // neighbour bit ^ bond bit ^ own bit
            uint32_t Y[6];
#pragma unroll
            for (int d = 0; d < D; ++d) Y[d] = A[d] ^ jb[d * B.jstride + jrow] ^ W;
            // unsatisfied-bond count m of every site: bits c0, c1, c2
            uint32_t c0, c1, c2;
            if (DIM == 3) {
                uint32_t s1, k1, s2, k2;
                fadd(Y[0], Y[1], Y[2], s1, k1);
                fadd(Y[3], Y[4], Y[5], s2, k2);
                c0 = s1 ^ s2;
                fadd(k1, k2, s1 & s2, c1, c2);
            } else {
                uint32_t s1, k1;
                fadd(Y[0], Y[1], Y[2], s1, k1);
                c0 = s1 ^ Y[3];
                const uint32_t k0 = s1 & Y[3];
                c1 = k1 ^ k0;
                c2 = k1 & k0;
            }
            const int par = DIM == 3 ? ((c + r + plane) & 1) : ((c + plane) & 1);
            // q = D - 2m <= 0 always flips: m >= 3 in 3-D, m >= 2 in 2-D.  Otherwise the
            // threshold of q: 3-D m = 2, 1, 0 -> t2, t4, t6; 2-D m = 1, 0 -> t2, t4.
            const uint32_t always = DIM == 3 ? (c2 | (c1 & c0)) : (c2 | c1);
            uint32_t flip = 0;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {          // the word's four segments
                pt_u4 rv;
                if (pre) {
                    const uint4 v = B.U[uo + (w << 2) + kk];
                    rv = pt_u4{v.x, v.y, v.z, v.w};
                } else {
                    const uint32_t q = ((uint32_t)(c * g.half) >> 2) +
                                       ((uint32_t)((plane << B.pb) + (r << Lb) + (wi << 5) +
                                                   (kk << 3)) >> 3);
                    rv = pt_philox(q, (uint32_t)a, t, PT_TAG_METRO, I.k0, I.k1);
                }
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) {
                    const uint32_t bit = 1u << (8 * kk + par + 2 * ii);
                    const uint32_t th = DIM == 3 ? ((c1 & bit) ? t2 : ((c0 & bit) ? t4 : t6))
                                                 : ((c0 & bit) ? t2 : t4);
                    flip |= (always & bit) | (pt_word(rv, ii) < th ? bit : 0u);
                }
            }
            sb[wrow] = W ^ flip;
            // energy change of the flips: sum of 2q = 2(D - 2m) over the flipped sites (in
            // signed arithmetic: __popc returns unsigned)
            if (owned) {
                const int nf = (int)__popc(flip);
                const int sm = (int)__popc(flip & c0) + 2 * (int)__popc(flip & c1) +
                               4 * (int)__popc(flip & c2);
                dsum += 2 * (D * nf - 2 * sm);
            }
        }
        __syncthreads();
    }
    return dsum;
}

// The first and last min(2, T) owned planes -- the neighbours' halos -- into nxt ([plane][word]).
__device__ __forceinline__ void bit_publish(const BitSlab& B, uint32_t* nxt) {
    const int nb = B.T < 2 ? B.T : 2;
    for (int idx = threadIdx.x; idx < 2 * nb * B.WPP; idx += blockDim.x) {
        const int kk = idx >> B.wppb, w = idx & (B.WPP - 1);
        const int o = kk < nb ? kk : B.T - 2 * nb + kk;
        st_u32(nxt + (size_t)(B.z0 + o) * B.WPP + w, B.sb[(size_t)(B.so + o) * B.WPP + w]);
    }
}

// The four halo slots (0, 1, T + 2, T + 3) from src ([plane][word]), entries first, first +
// stride, ...
__device__ __forceinline__ void bit_load_halos(const BitSlab& B, const uint32_t* src, int first,
                                               int stride) {
    for (int idx = first; idx < 4 * B.WPP; idx += stride) {
        const int kk = idx >> B.wppb, w = idx & (B.WPP - 1);
        const int sl = kk < 2 ? kk : B.T + kk;
        const int plane = (B.z0 - 2 + sl) & (B.L - 1);
        B.sb[(size_t)sl * B.WPP + w] = ld_u32(src + (size_t)plane * B.WPP + w);
    }
}

// bit_publish with each word tagged by the sweep count tag (high half): the reader polls the data
// itself, so no flag has to follow the stores.
__device__ __forceinline__ void bit_publish_tagged(const BitSlab& B, uint64_t* dst, uint32_t tag) {
    const int nb = B.T < 2 ? B.T : 2;
    for (int idx = threadIdx.x; idx < 2 * nb * B.WPP; idx += blockDim.x) {
        const int kk = idx >> B.wppb, w = idx & (B.WPP - 1);
        const int o = kk < nb ? kk : B.T - 2 * nb + kk;
        st_u64(dst + (size_t)(B.z0 + o) * B.WPP + w,
               ((uint64_t)tag << 32) | B.sb[(size_t)(B.so + o) * B.WPP + w]);
    }
}

// The four halo slots from the tagged words of src once each carries tag (one wave: entries
// lane, lane + 64, ...).  False when a wait exceeded its bound.
__device__ __forceinline__ bool bit_wait_halos(const BitSlab& B, const CoopArgs& C,
                                               const uint64_t* src, uint32_t tag, int lane) {
    for (int idx = lane; idx < 4 * B.WPP; idx += PT_WAVE) {
        const int kk = idx >> B.wppb, w = idx & (B.WPP - 1);
        const int sl = kk < 2 ? kk : B.T + kk;
        const int plane = (B.z0 - 2 + sl) & (B.L - 1);
        const uint64_t* p = src + (size_t)plane * B.WPP + w;
        const unsigned long long t0 = (unsigned long long)wall_clock64();
        uint64_t v;
        while ((uint32_t)((v = ld_u64(p)) >> 32) != tag) {
            if (__hip_atomic_load(C.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                (unsigned long long)wall_clock64() - t0 > C.timeout_ticks) {
                __hip_atomic_store(C.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        B.sb[(size_t)sl * B.WPP + w] = (uint32_t)v;
    }
    return true;
}

// At most 512 threads: the next-sweep uniform path needs more than 128 registers per lane.
// EBAR: the energy path of the barrier (SState::ebar), fixed per instantiation.
template <int DIM, bool WRAP, bool EBAR>
__global__ __launch_bounds__(512) void pslab_bits_run(PtInstance I, SState S, CoopArgs C) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lw[];
    __shared__ int64_t Es[1024];
    __shared__ int64_t EbL[1024];
    __shared__ int32_t reps[1024];
    __shared__ int32_t rungs[1024];
    __shared__ uint32_t thr3[3 * 1024];
    __shared__ long long scrE[16];
    __shared__ int scrR[16];
    __shared__ long long bestL[5];
    __shared__ double betaL[1024];
    __shared__ int32_t labelL[1024];
    __shared__ int sh_abort, sh_imp, sh_slot;
    __shared__ long long sh_dsum;
    __shared__ uint32_t sh_gen, sh_uready;
    __shared__ uint32_t sh_uex[PT_WAVE];
    __shared__ long long sh_esum[PT_WAVE];   // workgroup 0: slot sums of the arrival words
    int a, b;
    wg_slot_slab(blockIdx.x, S.NB, a, b);
    const bool live = a < I.R, owner = blockIdx.x == 0;
    const Geo g = make_geo<DIM>(I);
    const BitSlab B = make_bit_slab<DIM, WRAP>(g, S.NB, b, lw);
    const int L = g.L, P = g.P, tid = threadIdx.x, nt = blockDim.x;
    const int WPP = B.WPP, wppb = B.wppb, T = B.T, z0 = B.z0, z1 = z0 + T, so = B.so;
    const int R = I.R, Rp = I.Rp;
    uint32_t* sb = B.sb;
    uint32_t* hb = S.hbits;                               // [parity][slot][plane][word]
    const size_t hrep = (size_t)L * WPP;                  // words per replica and parity

    // ---- stage the slab (bits) and the bond planes ----------------------------------------------
    if (live) bit_stage<DIM, WRAP>(B, I, S, g, a);
    for (int x = tid; x < R; x += nt) {
        betaL[x] = I.beta[x];
        labelL[x] = (R > 1 && x == 0) ? 1 : ((R > 1 && x == R - 1) ? -1 : 0);
        const uint32_t* tr = I.thr + (size_t)x * I.TQ;
        thr3[3 * x] = I.TQ >= 2 ? tr[1] : 0u;
        thr3[3 * x + 1] = I.TQ >= 4 ? tr[3] : 0u;
        thr3[3 * x + 2] = (DIM == 3 && I.TQ >= 6) ? tr[5] : 0u;
    }
    if (tid == 0) {
        sh_abort = 0;
        sh_dsum = 0;
        sh_gen = 0;
        sh_uready = 0;
        for (int k = 0; k < 5; ++k) bestL[k] = 0;
    }
    if (tid < PT_WAVE) sh_esum[tid] = 0;
    // R <= 64: the round end is wave 0's (WaveLadder); the other waves meanwhile bring in the
    // halo planes of the round's last sweep
    const bool fast = R <= PT_WAVE;
    const bool defer = fast && !WRAP && nt > PT_WAVE && I.S > 0;
    WaveLadder WL;
    if (fast && tid < PT_WAVE) wave_ladder_init(WL, I, S.einit, tid);

    uint32_t phase = 0;
    // Next-sweep uniforms: with S.upre, waves 1.. draw the Philox blocks of sweep t + 1 into LDS
    // while wave 0 alone completes the grid barrier and the round end; the colour passes then read
    // them.
    const bool pre = S.upre != 0 && fast && nt > PT_WAVE;
    // ebar (with pre): the energy changes of a round reach wave 0 through the barrier itself.  Each
    // workgroup's arrival word carries its change; workgroup 0 sums them by slot into the release
    // words; lane a of every wave 0 keeps slot a's change (relv) for the round end.  Otherwise the
    // changes go through atomic adds into S.dacc before the arrival flags.
    const bool ebar = EBAR && pre;
    int64_t relv = 0;
    const uint32_t nsweeps = (uint32_t)I.n_rounds * (uint32_t)I.S;
    if (pre && live && nsweeps > 0) bit_draw<DIM, WRAP>(B, I, g, a, 0u, tid, nt);
    __syncthreads();
    for (int rnd = 0;; ++rnd) {
        // ---- end of round rnd - 1 (rnd == 0: the initial state), identical in every workgroup
        if (fast) {
            if (tid < PT_WAVE) {
                int64_t E = WL.Eb;
                if (ebar) {
                    const int64_t er = __shfl(relv, WL.rep & (PT_WAVE - 1));
                    if (rnd > 0 && tid < R) E += er;
                } else if (rnd > 0 && tid < R) {
                    E += ld_coh(S.dacc + (size_t)((rnd - 1) % 3) * Rp + WL.rep);
                }
                // pre: wave 1 drew this exchange's uniforms during the barrier (a short wait)
                if (pre && rnd > 0)
                    while (phase_before(__hip_atomic_load(&sh_uready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP), phase))
                        __builtin_amdgcn_s_sleep(1);
                const bool imp = wave_ladder_round(WL, I, rnd, E, tid, owner,
                                                   pre && rnd > 0 ? sh_uex : nullptr);
                if (tid < R) rungs[WL.rep] = tid;
                if (tid == 0) {
                    sh_imp = imp;
                    sh_slot = WL.best_slot;
                }
                if (owner && !ebar && tid < R)   // buffer of round rnd + 1 (see below)
                    __hip_atomic_store(S.dacc + (size_t)((rnd + 1) % 3) * Rp + tid, (int64_t)0,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (defer && live && rnd > 0 && rnd < I.n_rounds) {
                bit_load_halos(B, hb + ((size_t)(((uint32_t)rnd * (uint32_t)I.S) & 1) * Rp + a) * hrep,
                               tid - PT_WAVE, nt - PT_WAVE);
            }
            __syncthreads();
            if (sh_abort) return;
            if (sh_imp) {
                const int slot = sh_slot;
                if (live && slot == a)
                    for (int site = z0 * P + tid; site < z1 * P; site += nt) {
                        const int ls = site - z0 * P;
                        const uint32_t wd = sb[(size_t)so * WPP + (ls >> 5)];
                        st_spin<true>(I.best_s + pos_of<DIM>(site, g),
                                      ((wd >> (ls & 31)) & 1u) ? (int8_t)-1 : (int8_t)1);
                    }
                __syncthreads();   // the spins are read above before the sweep changes them
            }
            if (rnd == I.n_rounds) break;
        } else {
        for (int x = tid; x < R; x += nt) {
            if (rnd == 0) {
                Es[x] = -(S.einit[x] / 2);
                reps[x] = x;
                rungs[x] = x;
            } else {
                Es[x] = EbL[x] + ld_coh(S.dacc + (size_t)((rnd - 1) % 3) * Rp + x);
            }
        }
        __syncthreads();
        int64_t mE;
        int32_t mR;
        block_best_rung(Es, reps, R, scrE, scrR, mE, mR);
        const bool improved = bestL[4] == 0 || mE < bestL[0];
        const int slot = reps[mR];
        // best state of the owned planes, written through (see pslab_run)
        if (improved && live && slot == a)
            for (int site = z0 * P + tid; site < z1 * P; site += nt) {
                const int ls = site - z0 * P;
                const uint32_t wd = sb[(size_t)so * WPP + (ls >> 5)];
                st_spin<true>(I.best_s + pos_of<DIM>(site, g), ((wd >> (ls & 31)) & 1u) ? (int8_t)-1 : (int8_t)1);
            }
        if (owner && rnd > 0) {
            if (tid == 0 && I.hist) I.hist[rnd - 1] = (double)(improved ? mE : bestL[0]) * I.unit;
            if (I.trace)
                for (int r = tid; r < R; r += nt)
                    I.trace[(size_t)(rnd - 1) * R + r] = (double)Es[reps[r]] * I.unit;
        }
        __syncthreads();
        if (tid == 0 && improved) {
            bestL[0] = mE;
            bestL[1] = rnd;
            bestL[2] = mR;
            bestL[3] = slot;
            bestL[4] = 1;
        }
        if (rnd > 0) {
            exchange_pairs(I, betaL, rnd - 1, Es, reps, rungs,
                           owner ? pt_out_arr(I.out, R, PT_OUT_TRIED) : nullptr,
                           owner ? pt_out_arr(I.out, R, PT_OUT_ACC) : nullptr);
            __syncthreads();
            if (owner)
                update_labels(R, rungs, labelL, nullptr, pt_out_arr(I.out, R, PT_OUT_TRIPS),
                              pt_out_arr(I.out, R, PT_OUT_UP), pt_out_arr(I.out, R, PT_OUT_DOWN));
        }
        for (int x = tid; x < R; x += nt) EbL[x] = Es[x];
        if (owner)
            for (int x = tid; x < R; x += nt)
                __hip_atomic_store(S.dacc + (size_t)((rnd + 1) % 3) * Rp + x, (int64_t)0,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (rnd == I.n_rounds) break;
        }

        // ---- the sweeps of round rnd --------------------------------------------------------
        const int gr = live ? rungs[a] : 0;
        const uint32_t t2 = thr3[3 * gr], t4 = thr3[3 * gr + 1], t6 = thr3[3 * gr + 2];
        int64_t dsum = 0;
        for (int j = 0; j < I.S; ++j) {
            const uint32_t t = (uint32_t)rnd * (uint32_t)I.S + (uint32_t)j;
            uint32_t* nxt = hb + ((size_t)((t + 1) & 1) * Rp + a) * hrep;
            if (live) dsum += bit_sweep<DIM, WRAP>(B, I, g, a, t, pre, t2, t4, t6);
            if (live && !WRAP) bit_publish(B, nxt);
            const bool last = j == I.S - 1;
            if (last) {   // the slot's energy change: waves into LDS, thread 0 to HBM on arrival
                long long v = (long long)dsum;
#pragma unroll
                for (int m = PT_WAVE / 2; m > 0; m >>= 1) v += __shfl_xor(v, m);
                if ((tid & (PT_WAVE - 1)) == 0 && v != 0 && live)
                    atomicAdd((unsigned long long*)&sh_dsum, (unsigned long long)v);
                dsum = 0;
            }
            if (pre) {
                // arrival: every wave's stores complete, then wave 0 completes the barrier
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
                ++phase;
                if (tid < PT_WAVE && !ebar) {
                    // thread 0 adds the energy change of the round to the slot's accumulator and
                    // flags; the hub (one lane per flag) publishes the generation word
                    if (tid == 0) {
                        const long long v = sh_dsum;
                        sh_dsum = 0;
                        if (last && live && v != 0) {
                            atomicAdd((unsigned long long*)(S.dacc + (size_t)(rnd % 3) * Rp + a),
                                      (unsigned long long)v);
                            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                        }
                        if (blockIdx.x != 0) st_u32(C.flags + blockIdx.x, phase);
                    }
                    if (blockIdx.x == 0) {
                        for (int x = 1 + tid; x < (int)gridDim.x; x += PT_WAVE)
                            if (!wait_phase(C, C.flags + x, phase)) sh_abort = 1;
                        if (tid == 0 && sh_abort == 0) st_u32(C.gen, phase);
                    } else if (tid == 0) {
                        if (!wait_phase(C, C.gen, phase)) sh_abort = 1;
                    }
                    if (tid == 0)
                        __hip_atomic_store(&sh_gen, phase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else if (tid < PT_WAVE) {
                    long long v = 0;
                    if (tid == 0) {
                        v = last && live ? sh_dsum : 0;
                        sh_dsum = 0;
                        if (blockIdx.x != 0) st_u64(C.eflags + blockIdx.x, tag_word(phase, v));
                    }
                    if (blockIdx.x == 0) {
                        // the hub: lane x polls the words of workgroups x, x + 64, ... together
                        // (every pass loads each word, the loads in flight at once), adds each
                        // change to its slot's sum once its tag arrives; lane a then releases slot
                        // a's sum.  Workgroup 0 (slot 0) adds its own change directly.
                        const int W = (int)gridDim.x, nmax = (W + PT_WAVE - 1) / PT_WAVE;
                        uint32_t pending = 0;
                        for (int i = 0; i < nmax; ++i) {
                            const int x = tid + i * PT_WAVE;
                            if (x != 0 && x < W) pending |= 1u << i;
                        }
                        if (tid == 0 && v != 0)
                            atomicAdd((unsigned long long*)&sh_esum[0], (unsigned long long)v);
                        const unsigned long long t0 = (unsigned long long)wall_clock64();
                        const uint32_t tag = phase & 0xffffu;
                        while (pending) {
                            for (int i0 = 0; i0 < nmax; i0 += 8) {   // W <= 2048: nmax <= 32
                                uint64_t w[8];
#pragma unroll
                                for (int i = 0; i < 8; ++i)
                                    if (i0 + i < nmax)
                                        w[i] = ld_u64(C.eflags + min(tid + (i0 + i) * PT_WAVE, W - 1));
#pragma unroll
                                for (int i = 0; i < 8; ++i) {
                                    const int k = i0 + i;
                                    if (k < nmax && ((pending >> k) & 1u) &&
                                        (uint32_t)(w[i] >> 48) == tag) {
                                        pending &= ~(1u << k);
                                        const int64_t e = word_value(w[i]);
                                        int xa, xb;
                                        wg_slot_slab(tid + k * PT_WAVE, S.NB, xa, xb);
                                        if (e != 0 && xa < PT_WAVE)
                                            atomicAdd((unsigned long long*)&sh_esum[xa],
                                                      (unsigned long long)e);
                                    }
                                }
                            }
                            if (!pending) break;
                            if (__hip_atomic_load(C.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                                (unsigned long long)wall_clock64() - t0 > C.timeout_ticks) {
                                __hip_atomic_store(C.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                sh_abort = 1;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                        __builtin_amdgcn_wave_barrier();
                        relv = sh_esum[tid];
                        sh_esum[tid] = 0;
                        if (sh_abort == 0) {
                            st_u64(C.erel + tid, tag_word(phase, relv));
                            if (tid == 0) st_u32(C.gen, phase);
                        }
                    } else {
                        // the release is the one generation word; the sums follow in one load
                        // (their tags confirm them)
                        if (tid == 0 && !wait_phase(C, C.gen, phase)) sh_abort = 1;
                        uint64_t w;
                        if (!wait_tag(C, C.erel + tid, phase, w)) sh_abort = 1;
                        relv = word_value(w);
                    }
                    if (tid == 0)
                        __hip_atomic_store(&sh_gen, phase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    if (last && tid < 2 * PT_WAVE) {
                        // wave 1: the uniforms of this round's exchange, for wave 0's round end
                        const int r = tid - PT_WAVE;
                        if (r < R) {
                            const pt_u4 ev = pt_philox((uint32_t)(r >> 2), 0u, (uint32_t)rnd,
                                                       PT_TAG_EXCH, I.k0, I.k1);
                            sh_uex[r] = pt_word(ev, r & 3);
                        }
                        if (tid == PT_WAVE)
                            __hip_atomic_store(&sh_uready, phase, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    if (live && t + 1 < nsweeps) bit_draw<DIM, WRAP>(B, I, g, a, t + 1, tid - PT_WAVE, nt - PT_WAVE);
                    while (phase_before(__hip_atomic_load(&sh_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP), phase))
                        __builtin_amdgcn_s_sleep(1);
                }
                if (last) continue;   // the round end brings in the halo planes
                if (tid >= PT_WAVE && live && !WRAP) bit_load_halos(B, nxt, tid - PT_WAVE, nt - PT_WAVE);
                __syncthreads();
                if (sh_abort) return;
                continue;
            }
            if (!coop_sync(C, phase, &sh_abort, S.dacc + (size_t)(rnd % 3) * Rp + a,
                           last && live ? &sh_dsum : nullptr))
                return;
            if (last && defer) continue;   // halo planes come in during the round end
            if (live && !WRAP) bit_load_halos(B, nxt, tid, nt);
            __syncthreads();
        }
    }

    // ---- final state where the host reads it (int8 spins) ---------------------------------------
    if (live) {
        int8_t* fb = S.sp + ((size_t)(((int64_t)I.n_rounds * I.S) & 1) * Rp + a) * g.N;
        for (int idx = tid; idx < T * WPP; idx += nt) {
            const int o = idx >> wppb, rw = idx & (WPP - 1);
            bits_row(sb[(size_t)(so + o) * WPP + rw], fb + (size_t)(z0 + o) * P + ((size_t)rw << 5));
        }
    }
    if (owner && fast) {
        const int fin = I.n_rounds & 1;
        if (tid < PT_WAVE)
            wave_ladder_store(WL, I, S.Eb + (size_t)fin * Rp, S.rung + (size_t)fin * Rp,
                              S.rep + (size_t)fin * Rp, S.label, S.best + 5 * fin, tid);
    } else if (owner) {
        const int fin = I.n_rounds & 1;
        for (int x = tid; x < R; x += nt) {
            S.Eb[(size_t)fin * Rp + x] = EbL[x];
            S.rung[(size_t)fin * Rp + x] = rungs[x];
            S.rep[(size_t)fin * Rp + x] = reps[x];
            S.label[x] = labelL[x];
        }
        if (tid < 5) S.best[5 * fin + tid] = bestL[tid];
    }
}

// ---------------------------------------------------------------------------------------------
// Bit-sliced persistent slab engine with neighbour-scoped synchronisation.  The same sweeps as
// pslab_bits_run, but no workgroup waits for the whole grid:
//  * halos: a workgroup publishes the planes its neighbours need as words tagged with the sweep
//    count (NbrArgs::halo) and, before sweep t, polls the words of its four halo planes -- the
//    adjacent slabs of its own slot -- until they carry sweep t; one wave does this while wave 0
//    takes the round end and the others draw the next sweep's uniforms;
//  * round end: each workgroup posts its slab's energy change, tagged with the round, into the
//    mailbox of its slot's rung (NbrArgs::box; slab 0 adds the slot's energy of the last round
//    end, so the words of a rung sum to that slot's new energy).  Exchange partners read the two
//    rungs' words -- nothing else -- and take the same decision (the pair's uniform, betas and
//    energies); a slot carries its own rung, label and trip count.
//  * best: each slot keeps its own best (energy, then first round, at the rung it held) and its
//    state as bits (NbrArgs::sbest, bestbits); the host's lexicographic minimum over the slots is
//    the first round whose lowest rung reached the lowest energy, which the per-round scan of the
//    reference finds (cuda/ising_pt.cu:126-137 tracks per replica, 224-229 takes the minimum).
// Counters of rungs and pairs are added atomically (one writer per rung and round, but rounds of
// different slots overlap).  A workgroup is at most a few rounds ahead of any rung it reads (a rung
// pairs with each neighbour every second round), within the NBR_ROUNDS buffered rounds; a word
// whose tag is not the round awaited is not taken.  Best energy per round and the rung trace need
// the whole ladder every round: runs that record them use pslab_bits_run.
// ---------------------------------------------------------------------------------------------
template <int DIM, bool WRAP>
__global__ __launch_bounds__(512) void pslab_bits_nbr(PtInstance I, SState S, CoopArgs C,
                                                     NbrArgs N) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lw[];
    __shared__ uint32_t thr3[3 * 1024];
    __shared__ double betaL[1024];
    __shared__ int sh_abort, sh_imp, sh_rung;
    __shared__ long long sh_dsum;
    int a, b;
    wg_slot_slab(blockIdx.x, S.NB, a, b);
    if (a >= I.R) return;                                 // padding slots take no part
    const Geo g = make_geo<DIM>(I);
    const BitSlab B = make_bit_slab<DIM, WRAP>(g, S.NB, b, lw);
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & (PT_WAVE - 1);
    const int R = I.R, Rp = I.Rp, NB = S.NB;
    const size_t hrep = (size_t)g.L * B.WPP;              // words per replica
    // this slot's tagged plane words, two buffers by sweep parity: a neighbour can finish sweep
    // t + 1 (it needs only this workgroup's planes of sweep t) before this workgroup has read its
    // planes of sweep t, but not sweep t + 2
    uint64_t* halo = N.halo + (size_t)a * hrep;
    const size_t hpar = (size_t)Rp * hrep;

    bit_stage<DIM, WRAP>(B, I, S, g, a);
    for (int x = tid; x < R; x += nt) {
        betaL[x] = I.beta[x];
        const uint32_t* tr = I.thr + (size_t)x * I.TQ;
        thr3[3 * x] = I.TQ >= 2 ? tr[1] : 0u;
        thr3[3 * x + 1] = I.TQ >= 4 ? tr[3] : 0u;
        thr3[3 * x + 2] = (DIM == 3 && I.TQ >= 6) ? tr[5] : 0u;
    }
    if (tid == 0) {
        sh_abort = 0;
        sh_imp = 0;
        sh_rung = a;
        sh_dsum = 0;
    }
    const uint32_t nsweeps = (uint32_t)I.n_rounds * (uint32_t)I.S;
    // wave 0 takes the round end, wave 1 (wave 0 when alone) the halo planes, the waves after
    // (wave 1 too when there are two) draw the next sweep's uniforms
    const bool pre = S.upre != 0 && nt > PT_WAVE;
    const int hwave = nt > PT_WAVE ? 1 : 0;
    const int dfirst = nt > 2 * PT_WAVE ? 2 * PT_WAVE : PT_WAVE;
    if (pre && nsweeps > 0) bit_draw<DIM, WRAP>(B, I, g, a, 0u, tid, nt);
    // the slot's ladder state, the same in every workgroup of the slot (wave 0's copy is used)
    int64_t E = -(S.einit[a] / 2);
    int gr = a, trips = 0;
    int lab = (R > 1 && a == 0) ? 1 : ((R > 1 && a == R - 1) ? -1 : 0);
    int64_t bestE = E;
    int best_round = 0, best_rung = a;
    uint32_t* mybest = N.bestbits + (size_t)a * hrep;
    __syncthreads();
    // the initial state is every slot's first best (round 0)
    for (int idx = tid; idx < B.T * B.WPP; idx += nt)
        mybest[(size_t)B.z0 * B.WPP + idx] = B.sb[(size_t)B.so * B.WPP + idx];

    for (int rnd = 0; rnd < I.n_rounds; ++rnd) {
        const int gcur = sh_rung;
        const uint32_t t2 = thr3[3 * gcur], t4 = thr3[3 * gcur + 1], t6 = thr3[3 * gcur + 2];
        int64_t dsum = 0;
        for (int j = 0; j < I.S; ++j) {
            const uint32_t t = (uint32_t)rnd * (uint32_t)I.S + (uint32_t)j;
            const bool last = j == I.S - 1, more = t + 1 < nsweeps;
            dsum += bit_sweep<DIM, WRAP>(B, I, g, a, t, pre, t2, t4, t6);
            if (!WRAP && more) bit_publish_tagged(B, halo + ((t + 1) & 1) * hpar, t + 1);
            if (last) {
                long long v = (long long)dsum;
#pragma unroll
                for (int m = PT_WAVE / 2; m > 0; m >>= 1) v += __shfl_xor(v, m);
                if (lane == 0 && v != 0) atomicAdd((unsigned long long*)&sh_dsum, (unsigned long long)v);
                dsum = 0;
                __syncthreads();
            }
            if (tid < PT_WAVE && last) {
                // ---- round end: this slot and its exchange partner, from two rungs' mailboxes
                uint64_t* box = N.box + (size_t)(rnd & (NBR_ROUNDS - 1)) * Rp * NB;
                if (lane == 0)
                    st_u64(box + (size_t)gr * NB + b,
                           tag_word((uint32_t)rnd, (b == 0 ? E : 0) + (int64_t)sh_dsum));
                const int par = rnd & 1;
                const bool lo = R > 1 && gr >= par && ((gr - par) & 1) == 0 && gr + 1 < R;
                const bool hi = R > 1 && gr >= par + 1 && ((gr - 1 - par) & 1) == 0;
                const int pr = lo ? gr + 1 : (hi ? gr - 1 : -1);
                int64_t own = 0, oth = 0;
                bool bad = false;
                for (int x = lane; x < 2 * NB; x += PT_WAVE) {
                    const bool mine = x < NB;
                    if (!mine && pr < 0) continue;
                    uint64_t w = 0;
                    if (!wait_tag(C, box + (size_t)(mine ? gr : pr) * NB + (mine ? x : x - NB),
                                  (uint32_t)rnd, w)) {
                        bad = true;
                        break;
                    }
                    if (mine) own += word_value(w);
                    else oth += word_value(w);
                }
#pragma unroll
                for (int m = PT_WAVE / 2; m > 0; m >>= 1) {
                    own += __shfl_xor(own, m);
                    oth += __shfl_xor(oth, m);
                }
                if (__any(bad)) {
                    if (lane == 0) sh_abort = 1;
                } else {
                    const int64_t En = own;
                    const bool imp = En < bestE;
                    if (imp) {
                        bestE = En;
                        best_round = rnd + 1;
                        best_rung = gr;
                    }
                    if (R > 1) {
                        int64_t* o = I.out;
                        if (pr >= 0) {
                            const int rlo = lo ? gr : pr;
                            const pt_u4 rv = pt_philox((uint32_t)(rlo >> 2), 0u, (uint32_t)rnd,
                                                       PT_TAG_EXCH, I.k0, I.k1);
                            const int64_t Elo = lo ? En : oth, Ehi = lo ? oth : En;
                            const bool ok = pt_accept_exchange(betaL[rlo], betaL[rlo + 1],
                                                               (double)Elo * I.unit,
                                                               (double)Ehi * I.unit,
                                                               pt_word(rv, rlo & 3));
                            if (lo && b == 0 && lane == 0) {
                                bump(pt_out_arr(o, R, PT_OUT_TRIED) + gr);
                                if (ok) bump(pt_out_arr(o, R, PT_OUT_ACC) + gr);
                            }
                            if (ok) gr = pr;
                        }
                        if (gr == 0) {
                            if (lab == -1) ++trips;
                            lab = 1;
                        } else if (gr == R - 1) {
                            lab = -1;
                        }
                        if (b == 0 && lane == 0) {
                            if (lab == 1) bump(pt_out_arr(o, R, PT_OUT_UP) + gr);
                            else if (lab == -1) bump(pt_out_arr(o, R, PT_OUT_DOWN) + gr);
                        }
                    }
                    E = En;
                    if (lane == 0) {
                        sh_rung = gr;
                        sh_imp = imp;
                        sh_dsum = 0;
                    }
                }
            }
            // the halo planes of sweep t + 1, as their owners publish them
            if (!WRAP && more && tid / PT_WAVE == hwave)
                if (!bit_wait_halos(B, C, halo + ((t + 1) & 1) * hpar, t + 1, lane)) sh_abort = 1;
            if (pre && more && tid >= dfirst)
                bit_draw<DIM, WRAP>(B, I, g, a, t + 1, tid - dfirst, nt - dfirst);
            __syncthreads();
            if (sh_abort) return;
            if (last && sh_imp) {   // the owned planes after this round's sweeps: the slot's new best
                for (int idx = tid; idx < B.T * B.WPP; idx += nt)
                    mybest[(size_t)B.z0 * B.WPP + idx] = B.sb[(size_t)B.so * B.WPP + idx];
                __syncthreads();
            }
        }
    }

    // ---- final state and the slot's ladder bookkeeping --------------------------------------------
    int8_t* fb = S.sp + ((size_t)(((int64_t)I.n_rounds * I.S) & 1) * Rp + a) * g.N;
    for (int idx = tid; idx < B.T * B.WPP; idx += nt) {
        const int o = idx >> B.wppb, rw = idx & (B.WPP - 1);
        bits_row(B.sb[(size_t)(B.so + o) * B.WPP + rw],
                 fb + (size_t)(B.z0 + o) * g.P + ((size_t)rw << 5));
    }
    if (b == 0 && tid == 0) {
        const int fin = I.n_rounds & 1;
        S.Eb[(size_t)fin * Rp + a] = E;
        S.rung[(size_t)fin * Rp + a] = gr;
        S.rep[(size_t)fin * Rp + gr] = a;
        S.label[a] = lab;
        pt_out_arr(I.out, R, PT_OUT_TRIPS)[a] = trips;
        N.sbest[4 * (size_t)a] = bestE;
        N.sbest[4 * (size_t)a + 1] = best_round;
        N.sbest[4 * (size_t)a + 2] = best_rung;
        N.sbest[4 * (size_t)a + 3] = 1;
    }
}

template <int DIM, bool WRAP>
const void* pslab_fn() {
    return reinterpret_cast<const void*>(&pslab_run<DIM, WRAP>);
}

template <int DIM, bool WRAP>
const void* pslab_bits_fn(bool ebar) {
    return ebar ? reinterpret_cast<const void*>(&pslab_bits_run<DIM, WRAP, true>)
                : reinterpret_cast<const void*>(&pslab_bits_run<DIM, WRAP, false>);
}

template <int DIM, bool WRAP>
const void* pslab_nbr_fn() {
    return reinterpret_cast<const void*>(&pslab_bits_nbr<DIM, WRAP>);
}

const void* pslab_nbr_pick(int dim, bool wrap) {
    if (dim == 3) return wrap ? pslab_nbr_fn<3, true>() : pslab_nbr_fn<3, false>();
    return wrap ? pslab_nbr_fn<2, true>() : pslab_nbr_fn<2, false>();
}

const void* pslab_pick(int dim, bool wrap, bool bits, bool ebar = false) {
    if (bits) {
        if (dim == 3) return wrap ? pslab_bits_fn<3, true>(ebar) : pslab_bits_fn<3, false>(ebar);
        return wrap ? pslab_bits_fn<2, true>(ebar) : pslab_bits_fn<2, false>(ebar);
    }
    if (dim == 3) return wrap ? pslab_fn<3, true>() : pslab_fn<3, false>();
    return wrap ? pslab_fn<2, true>() : pslab_fn<2, false>();
}

}  // namespace

size_t pt_slab_lds_bytes(int n, int L, int dim, int NB) {
    (void)n;
    const size_t P = dim == 3 ? (size_t)L * L : (size_t)L;
    const size_t T = (size_t)((L + NB - 1) / NB);
    const size_t spins = NB == 1 ? (size_t)L : T + 4;     // staged spin planes
    const size_t signs = NB == 1 ? (size_t)L : T + 2;     // bond-sign planes of updated slots
    return ((spins + signs) * P + 15) & ~(size_t)15;
}

hipError_t pt_launch_slab_init(const PtInstance& I, const SState& S, hipStream_t s) {
    const int64_t total = (int64_t)I.R * I.n;
    slab_init_spins<<<(int)std::min<int64_t>((total + 255) / 256, 4096), 256, 0, s>>>(I, S);
    const dim3 grid((unsigned)std::min(64, (I.n + 255) / 256), (unsigned)I.R);
    if (I.ldim == 3) slab_init_energy<3><<<grid, 256, 0, s>>>(I, S);
    else slab_init_energy<2><<<grid, 256, 0, s>>>(I, S);
    return hipGetLastError();
}

hipError_t pt_launch_slab_sweep(const PtInstance& I, const SState& S, int j, int rl, int sweep,
                                int block, size_t lds, hipStream_t s) {
    const int grid = ((I.R + 7) / 8) * 8 * S.NB;
    const bool wrap = S.NB == 1;
    if (I.ldim == 3) {
        if (wrap) slab_sweep<3, true><<<grid, block, lds, s>>>(I, S, j, rl, sweep);
        else slab_sweep<3, false><<<grid, block, lds, s>>>(I, S, j, rl, sweep);
    } else {
        if (wrap) slab_sweep<2, true><<<grid, block, lds, s>>>(I, S, j, rl, sweep);
        else slab_sweep<2, false><<<grid, block, lds, s>>>(I, S, j, rl, sweep);
    }
    return hipGetLastError();
}

hipError_t pt_slab_set_lds(int dim, int NB, size_t lds) {
    const bool wrap = NB == 1;
    const void* fn = dim == 3 ? (wrap ? (const void*)slab_sweep<3, true> : (const void*)slab_sweep<3, false>)
                              : (wrap ? (const void*)slab_sweep<2, true> : (const void*)slab_sweep<2, false>);
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t pt_launch_slab_tick(const SState& S, int rounds, hipStream_t s) {
    slab_tick<<<1, 64, 0, s>>>(S.round_base, rounds);
    return hipGetLastError();
}

bool pt_pslab_bits(int L) { return L >= 32 && (L & 31) == 0; }

size_t pt_pslab_lds_bytes(int L, int dim, int NB, bool upre) {
    if (!pt_pslab_bits(L)) return pt_slab_lds_bytes(0, L, dim, NB);
    const size_t WPP = (dim == 3 ? (size_t)L : 1) * (size_t)(L / 32);
    const size_t T = (size_t)((L + NB - 1) / NB);
    const size_t spins = NB == 1 ? (size_t)L : T + 4;
    const size_t signs = NB == 1 ? (size_t)L : T + 2;
    const size_t words = (spins + 2 * (size_t)dim * signs) * WPP;
    if (!upre) return words * 4;
    // one Philox block (16 bytes) per 8-site segment of every slot a sweep updates
    const size_t segs = (dim == 3 ? (size_t)L : 1) * (size_t)(L / 8);
    const size_t slots = NB == 1 ? 2 * (size_t)L : 2 * T + 2;
    return ((words * 4 + 15) & ~(size_t)15) + slots * segs * 16;
}

// The smallest occupancy of the kernels a run may take (the two energy-path variants and, for bit
// slabs, the neighbour-synchronised kernel), so any of them can be launched on a grid this says is
// co-resident.
hipError_t pt_pslab_occupancy(int dim, int NB, int L, int block, size_t lds, int* per_cu) {
    *per_cu = 0;
    const int nv = pt_pslab_bits(L) ? 3 : 2;
    for (int v = 0; v < nv; ++v) {
        const void* f = v == 2 ? pslab_nbr_pick(dim, NB == 1) : pslab_pick(dim, NB == 1, pt_pslab_bits(L), v == 1);
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        int k = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, f, block, lds);
        if (e != hipSuccess) return e;
        *per_cu = v == 0 ? k : std::min(*per_cu, k);
    }
    return hipSuccess;
}

int pt_pslab_grid(const PtInstance& I, const SState& S) { return ((I.R + 7) / 8) * 8 * S.NB; }

hipError_t pt_launch_pslab(const PtInstance& I, const SState& S, const CoopArgs& C, int block,
                           size_t lds, bool cooperative, hipStream_t s) {
    PtInstance i = I;
    SState st = S;
    CoopArgs c = C;
    void* args[] = {(void*)&i, (void*)&st, (void*)&c};
    const void* f = pslab_pick(I.ldim, S.NB == 1, pt_pslab_bits(I.lL), S.ebar != 0);
    const dim3 grid(pt_pslab_grid(I, S));
    return cooperative ? hipLaunchCooperativeKernel(f, grid, dim3(block), args, (unsigned)lds, s)
                       : hipLaunchKernel(f, grid, dim3(block), args, lds, s);
}

hipError_t pt_launch_pslab_nbr(const PtInstance& I, const SState& S, const CoopArgs& C,
                               const NbrArgs& N, int block, size_t lds, bool cooperative,
                               hipStream_t s) {
    PtInstance i = I;
    SState st = S;
    CoopArgs c = C;
    NbrArgs nb = N;
    void* args[] = {(void*)&i, (void*)&st, (void*)&c, (void*)&nb};
    const void* f = pslab_nbr_pick(I.ldim, S.NB == 1);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const dim3 grid(pt_pslab_grid(I, S));
    return cooperative ? hipLaunchCooperativeKernel(f, grid, dim3(block), args, (unsigned)lds, s)
                       : hipLaunchKernel(f, grid, dim3(block), args, lds, s);
}
