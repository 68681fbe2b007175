// pt_engine.h -- device-side instance descriptors and the launchers of the gfx950 kernels.
// Internal to libdriftpt.so; the public surface is include/drift_pt.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Coupling modes (match drift_pt_weight_type).
enum : int { PT_WM_I8 = 1, PT_WM_I32 = 2, PT_WM_F32 = 3 };

// Per-instance statistics block in device memory (int64 entries):
//   [0] best units  [1] best round  [2] best rung  [3] best slot  [4] have-best flag
//   [5] rounds run (written by the engines that stop at a target)
//   then R-long arrays: E by slot, rung by slot, tried by rung, accepted by rung,
//   round trips by slot, up-visits by rung, down-visits by rung.
#define PT_OUT_HDR 6
enum : int { PT_OUT_E = 0, PT_OUT_RUNG, PT_OUT_TRIED, PT_OUT_ACC, PT_OUT_TRIPS, PT_OUT_UP, PT_OUT_DOWN, PT_OUT_NARR };
__host__ __device__ inline int64_t* pt_out_arr(int64_t* out, int R, int which) {
    return out + PT_OUT_HDR + (int64_t)which * R;
}
inline int64_t pt_out_len(int R) { return PT_OUT_HDR + (int64_t)PT_OUT_NARR * R; }

// One instance as the kernels see it.  Spins are relabelled by colour-list position p (colour c
// owns positions color_ptr[c] .. color_ptr[c+1]-1); CSR rows and columns use positions; orig[p]
// is the caller's spin index.  Device pointers throughout.
struct PtInstance {
    const int32_t* row_ptr;   // n + 1
    const void* edge;         // nnz edge records in row order: int8 weights -> uint32
                              // (position << 8 | weight byte); int32 / float -> uint2
                              // {position, weight bits}.  One load per neighbour.
    const void* ell;          // the first gq edge records of each row, row-major, zero records
                              // padding short rows and the rows up to a multiple of 4 positions;
                              // the neighbour field holds position * Rp (its spins' LDS offset) on
                              // the LDS engine, the position on the per-launch CSR engine; null on
                              // the other engines
    const void* h;            // n (position order) int32 / float, or null; zero-padded to a
                              // multiple of 4 positions where ell is set
    const int32_t* color_ptr; // k + 1
    const int32_t* orig;      // n
    const double* beta;       // R, rung 0 coldest
    const uint32_t* thr;      // R * TQ integer-mode thresholds
    int8_t* best_s;           // n, position order
    int8_t* final_s;          // R * n slot-major, position order, or null
    double* hist;             // n_rounds or null
    double* trace;            // n_rounds * R energies by rung, or null
    int64_t* out;             // statistics block
    int8_t* img;              // LDS engine: the initial spins as the LDS holds them ([position]
                              // [slot], n * Rp bytes rounded up to 16), written by the init launch
                              // and read by the launch that runs the rounds; null on other engines
    int64_t* e0;              // LDS engine: the initial energy of every slot (Rp), likewise
    const uint8_t* jbits;     // lattice engine: per site, bit d set when the bond in direction
                              // d (+x,-x,+y,-y,+z,-z) is -1; null otherwise
    double unit;              // energy per unit
    uint32_t k0, k1;          // Philox key (seed)
    int32_t n, k, R, Rp;      // Rp: replica slots padded (power of two <= 64, else multiple of 64)
    int32_t n_rounds, S, TQ, wm;
    int32_t ldim, lL;         // lattice dimension and side (0 when not a lattice)
    int32_t gq;               // edge records per row in ell (LDS engine 2, 3 or 4; per-launch
                              // CSR engine 4 or 8); longer rows finish from the CSR in chunks of
                              // the same size
    int32_t ell_rest;         // some row is longer than gq
    int32_t has_target;       // stop after the first round end whose best is <= target_units
    int64_t target_units;
    // Split engine: the slots are held by ng workgroups of rg consecutive slots each (the last
    // group may hold padding slots); img holds ng images of n * rg bytes (rounded up to 16), each
    // group's spins [position][slot].  box: [round parity][Rp slots][2] energy mailboxes, the low
    // and the high 32 bits of a slot's energy after a round, each word under the round's tag
    // (pt_split_word), then two round words (by parity) that group 0 tags once it has read every
    // slot's words; all bits set at the start.  Other engines: rg = Rp, ng = 1, box null.
    uint64_t* box;
    int32_t rg, ng;
    // Couplings in LDS (split engine, int8 couplings with |w| <= 7, n <= 4096): the CSR as 16-bit
    // records, position << 4 | (weight & 15), in row order; null otherwise.
    const uint16_t* crec;
    int32_t nnz;
    int32_t thr_lds;          // split engine, integer couplings: each round's threshold rows of the
                              // group's slots are copied into LDS (rg * TQ words fit 16 KB)
    // LDS and split engines, int8 couplings with rows longer than the first chunk (ell_rest): each
    // colour class's quads (cs/4 .. (ce-1)/4) listed longest row first, the classes one after
    // another, so the lanes of a wave take quads of like length and the remainder's passes, which
    // run until the wave's longest row ends, waste few slots; null otherwise (natural order)
    const int32_t* qorder;
    // Split engine with np > 1 parts per group: workgroup (group g, part q) holds the group's
    // whole state but sweeps only part q's share of each colour class -- the class's 32-position
    // words (word w = positions 32 w .. 32 w + 31) split into np runs, as equal as integer
    // division makes them, run q its own.  pbits: [sweep parity][Rp slots][nw words] spin words,
    // the step count (sweep t, colour c: t k + c + 1) over 32 spin bits (1 = -1), all zero at
    // the start.  pbox: two halves of
    // [round parity][np parts][Rp slots][2] words in box's format: each part's energy change on
    // traded classes since the start, then each part's copy of its group's slot energies (all
    // bits set at the start).  np = 1: null.
    // pflag: [ng groups][np parts] flags, PT_PFLAG_STRIDE words apart, each the last step whose
    // words the part has published (zero at the start).
    uint64_t* pbits;
    uint64_t* pbox;
    uint64_t* pflag;
    int32_t np, nw;
};
#define PT_PFLAG_STRIDE 8

// Split engine kernel arguments: the abort flag (set when a mailbox wait exceeds timeout_ticks;
// every workgroup then leaves), the bound of one wait in device wall-clock ticks, and the work
// queues: nq queues of items, queue q holding items[qoff[q] .. qoff[q+1]) with its own counter.
// A workgroup takes from queue blockIdx.x mod nq first (workgroups are dispatched to the eight
// XCDs round robin, so with nq = 8 an instance's groups share one XCD's L2) and then from the
// others.
struct SplitArgs {
    int32_t* abort;
    unsigned long long timeout_ticks;
    const int32_t* qoff;
    int32_t nq;
};

#define PT_DACC_SPLIT 32

// Global-engine state (replica-interleaved spins in HBM).  Two-deep buffers are indexed by
// round parity so the many workgroups of one launch only ever read what another launch wrote.
struct GState {
    int8_t* sp;      // n * Rp, [position][slot]
    int64_t* Eb;     // 2 * Rp: energy at the start of the round
    int64_t* dacc;   // 2 * PT_DACC_SPLIT * Rp: energy change accumulated during the round, split
                     // over PT_DACC_SPLIT counters (by block index) to spread atomic traffic
    int32_t* rung;   // 2 * Rp: slot -> rung
    int32_t* rep;    // 2 * Rp: rung -> slot
    int32_t* label;  // Rp: round-trip direction label (+1 last at cold end, -1 at hot end)
    int64_t* best;   // 2 * 5: E, round, rung, slot, have
    int32_t* round_base;  // round index that launch argument 'rl' is relative to; advanced on the
                          // device by pt_launch_g_tick so a captured chunk of rounds can replay
    uint32_t* spw;        // R <= 32 on the CSR engine: one word per position, bit a = slot a's spin
                          // (1 = -1); the colour steps and round ends use it instead of sp
    int32_t* stop;        // with PtInstance::has_target: 0 while running; r + 2 once the round end
                          // of round r (r = -1: the initial states) reached the target, after
                          // which every kernel of a later round returns at entry (starts at 0)
};

// Slab-engine state: a periodic lattice split into slabs of planes along its slowest axis; one
// workgroup owns one slab of one replica per sweep.  Spins are double-buffered by sweep parity,
// energy changes rotate through three accumulators by round, and the round-end bookkeeping of
// round r-1 runs at the start of round r's first sweep kernel.
struct SState {
    int8_t* sp;        // 2 * Rp * n: [buffer][slot][site]
    int64_t* Eb;       // 2 * Rp: energy at the start of round r, index r & 1
    int64_t* dacc;     // 3 * Rp: energy change of round r, index r % 3
    int64_t* einit;    // Rp: initial -2E (sum of s_i * local_i) per slot
    int32_t* rung;     // 2 * Rp: slot -> rung during round r, index r & 1
    int32_t* rep;      // 2 * Rp: rung -> slot
    int32_t* label;    // Rp
    int64_t* best;     // 2 * 5: best after the round-end scan of round r-1, index r & 1
    int32_t* round_base;
    uint32_t* hbits;   // bit-sliced persistent engine: [sweep parity][slot][plane][word] spin bits
                       // of the planes workgroups hand to their neighbours (1 = spin -1)
    int32_t NB;        // slabs per replica (1: the whole lattice wraps inside one workgroup)
    int32_t upre;      // bit-sliced persistent engine: the next sweep's Metropolis uniforms are
                       // drawn into LDS while the grid barrier and the round end are in progress
    int32_t ebar;      // with upre: the energy changes travel in the barrier words (CoopArgs
                       // eflags, erel) instead of atomic adds into dacc
};

// Persistent CSR engine (one cooperative launch).  Workgroup w of W owns ranks
// [cz w / W, cz (w+1) / W) of every colour class (cz = class size) for the whole run, with their
// row offsets and edge records copied into its LDS.  GState.dacc has three buffers here (round
// r uses r % 3).  Offsets are byte offsets into the dynamic LDS of one workgroup.
struct CoopArgs {
    uint32_t* flags;                // W arrival flags: the last barrier phase each workgroup
                                    // reached (start at 0)
    uint32_t* gen;                  // last phase every workgroup reached (starts at 0)
    uint64_t* eflags;               // bit-sliced persistent engine with next-sweep uniforms: W
                                    // arrival words, each the phase tag over the workgroup's
                                    // energy change of the round (start at 0)
    uint64_t* erel;                 // the same engine: 64 release words written by workgroup 0,
                                    // word a the phase tag over slot a's energy change (start at 0)
    int32_t* abort;                 // set when a barrier wait exceeded timeout_ticks
    unsigned long long timeout_ticks;  // wall_clock64 ticks one barrier may wait
    int32_t thr_lds;                // thresholds staged in LDS
    uint32_t off_cptr, off_lrp, off_edge, off_lsp, off_thr, off_beta, off_betaf, off_Eb, off_Es,
        off_red, off_rep, off_rung, off_label, off_stats, off_scr, off_sh;
};

hipError_t pt_launch_coop(const PtInstance& I, const GState& G, const CoopArgs& C, int block,
                          int grid, size_t lds, bool cooperative, hipStream_t s);
// Workgroups of the coop kernel one CU holds at this block size and LDS size.
hipError_t pt_coop_occupancy(int wm, bool uni, int block, size_t lds, int* per_cu);

size_t pt_slab_lds_bytes(int n, int L, int dim, int NB);
hipError_t pt_launch_slab_init(const PtInstance& I, const SState& S, hipStream_t s);
// Sweep j of round (*round_base + rl); j == 0 first finishes round rl-1.  sweep == 0 runs only
// that round end (the final one).
hipError_t pt_launch_slab_sweep(const PtInstance& I, const SState& S, int j, int rl, int sweep,
                                int block, size_t lds, hipStream_t s);
hipError_t pt_launch_slab_tick(const SState& S, int rounds, hipStream_t s);
hipError_t pt_slab_set_lds(int dim, int NB, size_t lds);
// Persistent slab engine: one cooperative launch of pt_pslab_grid workgroups runs every round
// (C.flags needs one entry per workgroup).  SState.dacc rotates through three buffers by round.
int pt_pslab_grid(const PtInstance& I, const SState& S);
// The bit-sliced variant runs when the side is a multiple of 32 (rows of 32-bit words).
bool pt_pslab_bits(int L);
size_t pt_pslab_lds_bytes(int L, int dim, int NB, bool upre = false);
hipError_t pt_pslab_occupancy(int dim, int NB, int L, int block, size_t lds, int* per_cu);
// Neighbour-synchronised bit-sliced persistent engine (no grid barrier): round-tagged energy
// mailboxes by rung, NBR_ROUNDS rounds deep, sweep-tagged halo words, per-slot best records.
enum : int { NBR_ROUNDS = 8 };
struct NbrArgs {
    uint64_t* box;        // [NBR_ROUNDS][Rp][NB]: round tag over a slab's energy word (all bits
                          // set at the start: a tag no round of the first NBR_ROUNDS carries)
    uint64_t* halo;       // [sweep parity][Rp][plane][word]: sweep count over the spin bits of
                          // each plane a workgroup hands to its neighbours (all bits set at the
                          // start)
    int64_t* sbest;       // [Rp][4]: each slot's best energy, its round, the rung it held, 1
    uint32_t* bestbits;   // [Rp][plane][word]: each slot's best state as spin bits (1 = -1)
};
hipError_t pt_launch_pslab_nbr(const PtInstance& I, const SState& S, const CoopArgs& C,
                               const NbrArgs& N, int block, size_t lds, bool cooperative,
                               hipStream_t s);
hipError_t pt_launch_pslab(const PtInstance& I, const SState& S, const CoopArgs& C, int block,
                           size_t lds, bool cooperative, hipStream_t s);

// LDS engine: dynamic LDS of one workgroup (spins of all Rp slots, ladder arrays, the k + 1
// colour-class bounds).
size_t pt_lds_bytes(int n, int Rp, int k);

// Error reporting shared by every entry point of the library (drift_pt_last_error).
int pt_fail(int code, const char* msg);
void pt_clear_error();

// One persistent launch of the LDS engine over the work queue *d_queue (start at 0).  init: each
// workgroup draws its instances' initial spins and energies into PtInstance::img / e0; otherwise
// it loads them from there and runs every round.
hipError_t pt_launch_lds(const PtInstance* d_insts, const int32_t* d_order, int n_inst,
                         int32_t* d_queue, int wm, int block, int grid, size_t lds_bytes,
                         bool init, hipStream_t s);
// Bytes of PtInstance::img for n positions and Rp slots.
__host__ __device__ inline size_t pt_lds_img_bytes(int n, int Rp) { return ((size_t)n * Rp + 15) & ~(size_t)15; }

// Split engine: dynamic LDS of one workgroup holding rg of an instance's Rp slots (its spins,
// and the ladder arrays of all Rp slots) and, with couplings in LDS, the row pointers, the 16-bit
// records and the field.
// thr_words: the group's threshold rows staged in LDS (rg * TQ, or 0).
// parts: the extra arrays of a workgroup of a group split into parts (initial energies of every
// slot, this part's energy change of each of its slots).
size_t pt_split_lds_bytes(int n, int rg, int Rp, int k, int thr_words, bool cl = false, int nnz = 0,
                          bool has_h = false, bool parts = false);
// One persistent launch of the split engine over work items items[0 .. n_items) (instance index
// << 10 | group; with parts, instance << 10 | group * np + part; the init launch always takes
// groups), taken from the queue *d_queue (start at 0), the items of one instance consecutive.
// init: each item draws its group's initial spins and energies (no waiting); otherwise the groups
// of an instance run every round together, so every group of the largest instance must be
// resident at once: grid <= the co-resident capacity (pt_split_occupancy), cooperative unless
// plain.
// cl: every instance's couplings staged in LDS (PtInstance::crec; int8 couplings only).
// parts: the kernel of groups split into parts (PtInstance::np, any np >= 1; not with cl).
hipError_t pt_launch_split(const PtInstance* d_insts, const int32_t* d_items, int n_items,
                           int32_t* d_queue, int wm, int block, int grid, size_t lds_bytes,
                           bool init, const SplitArgs& A, bool cooperative, hipStream_t s,
                           bool cl = false, bool parts = false);
hipError_t pt_split_occupancy(int wm, int block, size_t lds_bytes, int* per_cu, bool cl = false,
                              bool parts = false);

hipError_t pt_launch_g_init(const PtInstance& I, const GState& G, int block, int round_end_grid,
                            hipStream_t s);
// Colour step of sweep j of round (*round_base + rl) over colour-list positions [cs, ce) (one
// colour class; passed by value so the first loads of the step do not wait on the colour list).
hipError_t pt_launch_g_color(const PtInstance& I, const GState& G, int cs, int ce, int j, int rl,
                             int block, int grid, hipStream_t s);
hipError_t pt_launch_g_lattice(const PtInstance& I, const GState& G, int c, int j, int rl,
                               int block, int grid, hipStream_t s);
// Word-per-position spins (R <= 32): pack sp into spw, and back.
hipError_t pt_launch_g_pack(const PtInstance& I, const GState& G, bool unpack, hipStream_t s);
// End of round (*round_base + rl): energies, best state, exchange.
hipError_t pt_launch_g_round_end(const PtInstance& I, const GState& G, int rl,
                                 int round_end_grid, hipStream_t s);
// *round_base += rounds.
hipError_t pt_launch_g_tick(const GState& G, int rounds, hipStream_t s);
