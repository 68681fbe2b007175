// pt_api.hip -- the C ABI of libdriftpt.so (include/drift_pt.h): validation, colouring,
// relabelling by colour position, threshold tables, device buffers, launches and result copies.
//
// Reference counterpart: the host half of cuda/ising_pt.cu (main, :159-243) plus the Python glue
// that fed it through a file and a subprocess (drift/gpu.py:59-114, CSR :66-77, colouring :122-145).
#include "../../include/drift_pt.h"
#include "pt_engine.h"
#include "pt_rules.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

int pt_fail(int code, const char* msg) { return fail(code, msg); }
void pt_clear_error() { g_err.clear(); }

namespace {

#define PT_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? DRIFT_PT_ENOMEM : DRIFT_PT_EHIP,          \
                        std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Device memory of one call.  By default every allocation is a hipMalloc, freed on every exit
// path.  With a fixed region (the allocation-free batch entry point) allocations are 256-byte
// aligned slices of that region instead: nothing is allocated or freed, and a request past its end
// fails with hipErrorOutOfMemory, leaving in `need` the bytes the call asked for up to then.
struct DevArena {
    std::vector<void*> ptrs;
    uint8_t* region = nullptr;
    size_t cap = 0, used = 0, need = 0;
    ~DevArena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** out, size_t count) {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (region) {
            const size_t at = (used + 255) & ~(size_t)255;
            need = at + bytes;
            if (need > cap) {
                *out = nullptr;
                return hipErrorOutOfMemory;
            }
            used = need;
            *out = (T*)(region + at);
            return hipSuccess;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) ptrs.push_back(p);
        *out = (T*)p;
        return e;
    }
    template <class T>
    hipError_t upload(T** out, const std::vector<T>& v) {
        hipError_t e = alloc(out, v.size());
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    }
};

// Device events for per-launch timing, destroyed with the call.
struct EventPool {
    std::vector<hipEvent_t> ev;
    ~EventPool() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    hipError_t pair(hipEvent_t* a, hipEvent_t* b) {
        hipError_t e = hipEventCreate(a);
        if (e != hipSuccess) return e;
        ev.push_back(*a);
        e = hipEventCreate(b);
        if (e != hipSuccess) return e;
        ev.push_back(*b);
        return hipSuccess;
    }
};

struct GraphGuard {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphGuard() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~StreamGuard() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

// One instance prepared on the host: relabelled by colour position, thresholds built.
struct HostInst {
    int n = 0, nnz = 0, k = 0, R = 0, Rp = 0, wm = 0, TQ = 1, S = 0, n_rounds = 0, flags = 0;
    uint64_t seed = 0;
    double unit = 1.0;
    std::vector<int32_t> row_ptr, color_ptr, orig;
    std::vector<uint32_t> edge;     // device edge records (see PtInstance::edge)
    std::vector<int32_t> h32;
    std::vector<float> hf;
    bool has_h = false;
    std::vector<double> beta;
    std::vector<uint32_t> thr;
    int ldim = 0, lL = 0;           // periodic hypercubic lattice detected (see detect_lattice)
    std::vector<uint8_t> jbits;     // per site: bit d set when the bond in direction d is -1
    int gq = 4;                     // PtInstance::gq
    int maxdeg = 0;                 // longest row
    int force_ebar = -1;            // bit-sliced persistent engine: SState::ebar, or -1 to choose
                                    // by a timed trial (energy_path_choice)
    int has_target = 0;             // DRIFT_PT_STOP_AT_TARGET: stop at best <= target_units
    int64_t target_units = 0;
    int rg_req = 0;                 // drift_pt_params::slots_per_group (0: the library chooses)
    int np_req = 0;                 // drift_pt_params::parts_per_group (0: the library chooses)
    bool cl_ok = false;             // couplings fit 16-bit LDS records (int8, |w| <= 7, n <= 4096)
    int block_req = 0;              // drift_pt_params::block_threads (0: the library chooses)
};

// First-chunk width of the LDS and split engines' quad gather: the smallest of 2, 3, 4 (and 6 for
// int8 couplings, whose records are one word) that covers at least three rows in four, else the
// widest.  A wider chunk costs issue slots on every shorter row; a row longer than the chunk costs
// one more wait on memory.  Wider chunks would hold more than 128 registers per lane (four rows
// of edge records and spins) and spill: width 8 on int8 records spilled 40 bytes per lane.
int quad_chunk(const std::vector<int32_t>& row_ptr, int n, int wm) {
    const int top = wm == PT_WM_I8 ? 6 : 4;
    int cnt[7] = {0};
    for (int p = 0; p < n; ++p) cnt[std::min(row_ptr[p + 1] - row_ptr[p], top)]++;
    int covered = 0;
    for (int d = 0; d <= top; ++d) {
        covered += cnt[d];
        if (d >= 2 && d != 5 && 4 * (int64_t)covered >= 3 * (int64_t)n) return d;
    }
    return top;
}

int32_t greedy_color(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* cptr,
                     int32_t* cspins) {
    std::vector<int32_t> color((size_t)std::max(n, 0), -1);
    int32_t maxdeg = 0;
    for (int32_t i = 0; i < n; ++i) maxdeg = std::max(maxdeg, rp[i + 1] - rp[i]);
    std::vector<int32_t> stamp((size_t)maxdeg + 2, -1);
    int32_t k = 1;
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t t = rp[i]; t < rp[i + 1]; ++t) {
            const int32_t c = color[ci[t]];
            if (c >= 0 && c <= maxdeg) stamp[c] = i;
        }
        int32_t c = 0;
        while (stamp[c] == i) ++c;
        color[i] = c;
        k = std::max(k, c + 1);
    }
    std::fill(cptr, cptr + k + 1, 0);
    for (int32_t i = 0; i < n; ++i) cptr[color[i] + 1] += 1;
    for (int32_t c = 0; c < k; ++c) cptr[c + 1] += cptr[c];
    std::vector<int32_t> fill(cptr, cptr + k);
    for (int32_t i = 0; i < n; ++i) cspins[fill[color[i]]++] = i;
    return k;
}

int padded_slots(int R) {
    if (R > 64) return ((R + 63) / 64) * 64;
    int p = 1;
    while (p < R) p <<= 1;
    return p;
}

int prepare(const drift_pt_problem* P, const drift_pt_params* prm, HostInst& H) {
    if (!P || !prm) return fail(DRIFT_PT_EINVAL, "null problem or params");
    const int n = P->n, nnz = P->nnz;
    if (n < 1) return fail(DRIFT_PT_EINVAL, "n must be >= 1");
    if (nnz < 0 || (nnz > 0 && (!P->col_idx || !P->weight)) || !P->row_ptr)
        return fail(DRIFT_PT_EINVAL, "missing CSR arrays");
    if (P->weight_type != DRIFT_PT_W_I8 && P->weight_type != DRIFT_PT_W_I32 &&
        P->weight_type != DRIFT_PT_W_F32)
        return fail(DRIFT_PT_EINVAL, "unknown weight_type");
    const int R = prm->n_replicas;
    if (R < 1 || R > 1024) return fail(DRIFT_PT_EINVAL, "n_replicas must be in [1, 1024]");
    if (prm->n_rounds < 0 || prm->sweeps_per_round < 0)
        return fail(DRIFT_PT_EINVAL, "n_rounds and sweeps_per_round must be >= 0");
    if ((int64_t)prm->n_rounds * std::max(prm->sweeps_per_round, 1) >= ((int64_t)1 << 32))
        return fail(DRIFT_PT_ERANGE, "more than 2^32 sweeps");
    if (prm->slots_per_group < 0 || prm->slots_per_group > 64 ||
        (prm->slots_per_group & (prm->slots_per_group - 1)) != 0)
        return fail(DRIFT_PT_EINVAL, "slots_per_group must be 0 or a power of two up to 64");
    if (prm->parts_per_group < 0 || prm->parts_per_group > 64)
        return fail(DRIFT_PT_EINVAL, "parts_per_group must be in [0, 64]");
    const int32_t* rp = P->row_ptr;
    const int32_t* ci = P->col_idx;
    if (rp[0] != 0 || rp[n] != nnz) return fail(DRIFT_PT_EINVAL, "row_ptr must run 0..nnz");
    for (int i = 0; i < n; ++i) {
        if (rp[i + 1] < rp[i]) return fail(DRIFT_PT_EINVAL, "row_ptr not monotone");
        for (int t = rp[i]; t < rp[i + 1]; ++t) {
            if (ci[t] < 0 || ci[t] >= n) return fail(DRIFT_PT_EINVAL, "col_idx out of range");
            if (ci[t] == i) return fail(DRIFT_PT_EINVAL, "self coupling on the diagonal");
        }
    }
    const bool isf = P->weight_type == DRIFT_PT_W_F32;
    auto wval = [&](int t) -> double {
        if (P->weight_type == DRIFT_PT_W_I8) return (double)((const int8_t*)P->weight)[t];
        if (P->weight_type == DRIFT_PT_W_I32) return (double)((const int32_t*)P->weight)[t];
        return (double)((const float*)P->weight)[t];
    };
    auto hval = [&](int i) -> double {
        if (!P->h) return 0.0;
        return isf ? (double)((const float*)P->h)[i] : (double)((const int32_t*)P->h)[i];
    };
    if (!(prm->flags & DRIFT_PT_SKIP_VALIDATION)) {
        // symmetry: every (i -> j, w) has its (j -> i, w); no duplicate columns within a row
        std::vector<std::pair<int32_t, double>> row;
        std::vector<int32_t> sorted_cols(ci, ci + nnz);
        for (int i = 0; i < n; ++i) std::sort(sorted_cols.begin() + rp[i], sorted_cols.begin() + rp[i + 1]);
        for (int i = 0; i < n; ++i)
            for (int t = rp[i]; t + 1 < rp[i + 1]; ++t)
                if (sorted_cols[t] == sorted_cols[t + 1])
                    return fail(DRIFT_PT_EINVAL, "duplicate column in a CSR row");
        for (int i = 0; i < n; ++i)
            for (int t = rp[i]; t < rp[i + 1]; ++t) {
                const int j = ci[t];
                bool found = false;
                for (int u = rp[j]; u < rp[j + 1]; ++u)
                    if (ci[u] == i) {
                        found = wval(u) == wval(t);
                        break;
                    }
                if (!found) return fail(DRIFT_PT_EINVAL, "coupling matrix is not symmetric");
            }
    }

    // colouring: given or greedy
    std::vector<int32_t> cptr, cspins((size_t)n);
    int k = P->n_colors;
    if (k <= 0) {
        cptr.assign((size_t)n + 1, 0);
        k = greedy_color(n, rp, ci, cptr.data(), cspins.data());
        cptr.resize((size_t)k + 1);
    } else {
        if (!P->color_ptr || !P->color_spins) return fail(DRIFT_PT_ECOLOR, "missing colour arrays");
        cptr.assign(P->color_ptr, P->color_ptr + k + 1);
        std::copy(P->color_spins, P->color_spins + n, cspins.begin());
        if (cptr[0] != 0 || cptr[k] != n) return fail(DRIFT_PT_ECOLOR, "color_ptr must run 0..n");
        for (int c = 0; c < k; ++c)
            if (cptr[c + 1] < cptr[c]) return fail(DRIFT_PT_ECOLOR, "color_ptr not monotone");
    }
    std::vector<int32_t> pos((size_t)n, -1), colour_of((size_t)n, -1);
    for (int c = 0; c < k; ++c)
        for (int p = cptr[c]; p < cptr[c + 1]; ++p) {
            const int i = cspins[p];
            if (i < 0 || i >= n || pos[i] >= 0)
                return fail(DRIFT_PT_ECOLOR, "color_spins is not a permutation of 0..n-1");
            pos[i] = p;
            colour_of[i] = c;
        }
    for (int i = 0; i < n; ++i)
        for (int t = rp[i]; t < rp[i + 1]; ++t)
            if (colour_of[ci[t]] == colour_of[i])
                return fail(DRIFT_PT_ECOLOR, "two adjacent spins share a colour");

    H = HostInst();
    H.n = n; H.nnz = nnz; H.k = k; H.R = R; H.Rp = padded_slots(R);
    H.wm = P->weight_type; H.S = prm->sweeps_per_round; H.n_rounds = prm->n_rounds;
    H.flags = prm->flags; H.seed = prm->seed; H.has_h = P->h != nullptr;
    H.rg_req = prm->slots_per_group;
    H.np_req = prm->parts_per_group;
    H.block_req = prm->block_threads;
    H.color_ptr = cptr;
    H.orig = cspins;
    H.row_ptr.assign((size_t)n + 1, 0);
    for (int p = 0; p < n; ++p) H.row_ptr[p + 1] = H.row_ptr[p] + (rp[cspins[p] + 1] - rp[cspins[p]]);
    // int8 weights pack with a 24-bit position into one word; larger instances use two words
    if (H.wm == PT_WM_I8 && n >= (1 << 24)) H.wm = PT_WM_I32;
    const int ew = H.wm == PT_WM_I8 ? 1 : 2;
    H.edge.resize((size_t)nnz * ew);
    if (H.has_h) {
        if (isf) H.hf.resize((size_t)n);
        else H.h32.resize((size_t)n);
    }
    double maxfield = 0.0, total_abs = 0.0;
    for (int p = 0; p < n; ++p) {
        const int i = cspins[p];
        int dst = H.row_ptr[p];
        double field = std::fabs(hval(i));
        for (int t = rp[i]; t < rp[i + 1]; ++t, ++dst) {   // row order kept: fixed summation order
            const uint32_t c = (uint32_t)pos[ci[t]];
            uint32_t wb;
            if (P->weight_type == DRIFT_PT_W_I8) wb = (uint32_t)(int32_t)((const int8_t*)P->weight)[t];
            else if (P->weight_type == DRIFT_PT_W_I32) wb = (uint32_t)((const int32_t*)P->weight)[t];
            else std::memcpy(&wb, (const float*)P->weight + t, 4);
            if (ew == 1) {
                H.edge[dst] = (c << 8) | (wb & 0xffu);
            } else {
                H.edge[2 * (size_t)dst] = c;
                H.edge[2 * (size_t)dst + 1] = wb;
            }
            field += std::fabs(wval(t));
        }
        if (H.has_h) {
            if (isf) H.hf[p] = ((const float*)P->h)[i];
            else H.h32[p] = ((const int32_t*)P->h)[i];
        }
        maxfield = std::max(maxfield, field);
        total_abs += field;
    }
    if (!std::isfinite(maxfield)) return fail(DRIFT_PT_ERANGE, "non-finite coupling or field");
    H.gq = quad_chunk(H.row_ptr, n, H.wm);
    if (H.wm == PT_WM_I8 && n <= 4096) {
        H.cl_ok = true;
        for (size_t t = 0; t < H.edge.size() && H.cl_ok; ++t) {
            const int w = (int)(int8_t)(H.edge[t] & 0xffu);
            H.cl_ok = w >= -7 && w <= 7;
        }
    }
    for (int p = 0; p < n; ++p) H.maxdeg = std::max(H.maxdeg, H.row_ptr[p + 1] - H.row_ptr[p]);
    if (isf) {
        if (maxfield >= 16384.0 || total_abs >= 1.0e8)
            return fail(DRIFT_PT_ERANGE, "float couplings too large for the 2^-36 fixed-point energy "
                                         "(local fields up to 2^14, energies up to 1e8)");
        H.unit = PT_FIX36_UNIT;
    } else {
        if (maxfield >= (double)(1 << 30)) return fail(DRIFT_PT_ERANGE, "local field exceeds 2^30");
        if (!(P->energy_scale > 0.0) || !std::isfinite(P->energy_scale))
            return fail(DRIFT_PT_EINVAL, "energy_scale must be positive");
        H.unit = P->energy_scale;
    }

    if (prm->flags & DRIFT_PT_STOP_AT_TARGET) {
        const double x = prm->target_energy / H.unit;
        if (std::isnan(x)) return fail(DRIFT_PT_EINVAL, "target_energy is NaN");
        H.has_target = 1;
        // beyond the units' range a target is never reached (below) or reached at once (above)
        H.target_units = x <= -9.0e18 ? INT64_MIN
                       : x >= 9.0e18  ? INT64_MAX
                                      : (int64_t)std::floor(x + 1e-9 * std::max(1.0, std::fabs(x)));
    }

    // ladder (drift/solvers/parallel_tempering.py:29-37, cuda/ising_pt.cu:174-178)
    H.beta.resize((size_t)R);
    if (prm->beta) {
        for (int r = 0; r < R; ++r) {
            if (!(prm->beta[r] > 0.0) || !std::isfinite(prm->beta[r]))
                return fail(DRIFT_PT_EINVAL, "beta entries must be positive and finite");
            H.beta[r] = prm->beta[r];
        }
    } else {
        if (!(prm->T_min > 0.0) || !(prm->T_max > 0.0))
            return fail(DRIFT_PT_EINVAL, "T_min and T_max must be positive");
        for (int r = 0; r < R; ++r) {
            const double T = R == 1 ? prm->T_min
                                    : prm->T_min * std::pow(prm->T_max / prm->T_min,
                                                            (double)r / (double)(R - 1));
            H.beta[r] = 1.0 / std::max(T, 1e-12);
        }
    }
    // the float decision computes -(beta (2 q)) as (-2 beta) q, the same float when 2 beta is finite
    if (isf)
        for (double b : H.beta)
            if (b >= 0x1p100)
                return fail(DRIFT_PT_ERANGE, "beta entries must be below 2^100 with float couplings");

    // integer thresholds [rung][q-1]; moves with q > TQ are rejected at every rung
    if (!isf) {
        const int qmax = std::max(1, (int)maxfield);
        double bmin = H.beta[0];
        for (double b : H.beta) bmin = std::min(bmin, b);
        int tq = 1;
        while (tq < qmax && pt_threshold_int(bmin * H.unit, tq + 1) != 0u) ++tq;
        if ((int64_t)tq * R > ((int64_t)1 << 24))
            return fail(DRIFT_PT_ERANGE, "threshold table too large (tiny beta with huge couplings)");
        H.TQ = tq;
        H.thr.resize((size_t)R * tq);
        for (int r = 0; r < R; ++r)
            for (int q = 1; q <= tq; ++q) H.thr[(size_t)r * tq + q - 1] = pt_threshold_int(H.beta[r] * H.unit, q);
    }
    return 0;
}

// Recognise a periodic L^dim lattice (dim 3 or 2, L even >= 4) with +/-1 couplings, no field and
// the parity colouring in index order (what greedy colouring gives such a lattice).  Fills the
// per-site sign byte used by the lattice kernel.  Directions: 0 +x, 1 -x, 2 +y, 3 -y, 4 +z, 5 -z.
void detect_lattice(const drift_pt_problem* P, HostInst& H) {
    H.ldim = 0;
    if (P->weight_type != DRIFT_PT_W_I8 || H.has_h || H.k != 2 || (H.n & 1)) return;
    const int n = H.n;
    for (int dim = 3; dim >= 2; --dim) {
        const int L = (int)std::lround(std::pow((double)n, 1.0 / dim));
        int64_t nn = 1;
        for (int d = 0; d < dim; ++d) nn *= L;
        if (nn != n || L < 4 || (L & 1) || P->nnz != 2 * dim * n) continue;
        std::vector<uint8_t> jb((size_t)n, 0);
        const int32_t* rp = P->row_ptr;
        const int32_t* ci = P->col_idx;
        const int8_t* w = (const int8_t*)P->weight;
        bool ok = true;
        for (int i = 0; i < n && ok; ++i) {
            const int x = i % L, y = (i / L) % L, z = dim == 3 ? i / (L * L) : 0;
            int nb[6];
            auto idx = [&](int xx, int yy, int zz) { return (zz * L + yy) * L + xx; };
            nb[0] = idx((x + 1) % L, y, z); nb[1] = idx((x + L - 1) % L, y, z);
            nb[2] = idx(x, (y + 1) % L, z); nb[3] = idx(x, (y + L - 1) % L, z);
            if (dim == 3) { nb[4] = idx(x, y, (z + 1) % L); nb[5] = idx(x, y, (z + L - 1) % L); }
            if (rp[i + 1] - rp[i] != 2 * dim) { ok = false; break; }
            unsigned seen = 0;
            for (int t = rp[i]; t < rp[i + 1]; ++t) {
                int d = 0;
                while (d < 2 * dim && nb[d] != ci[t]) ++d;
                if (d == 2 * dim || (seen >> d & 1u) || (w[t] != 1 && w[t] != -1)) { ok = false; break; }
                seen |= 1u << d;
                if (w[t] == -1) jb[i] |= (uint8_t)(1u << d);
            }
        }
        if (!ok || H.color_ptr[1] != n / 2) continue;
        // parity colouring in index order
        int p0 = 0, p1 = n / 2;
        for (int i = 0; i < n && ok; ++i) {
            const int x = i % L, y = (i / L) % L, z = dim == 3 ? i / (L * L) : 0;
            const int par = (x + y + z) & 1;
            if (H.orig[par ? p1++ : p0++] != i) ok = false;
        }
        if (!ok) continue;
        H.ldim = dim;
        H.lL = L;
        H.jbits.swap(jb);
        return;
    }
}

int device_lds_limit() {
    static int cached = -1;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (cached < 0) {
        hipDeviceProp_t prop;
        int dev = 0;
        cached = 65536;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cached = (int)prop.sharedMemPerBlock;
    }
    return cached;
}

int check_device() {
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) {
        (void)hipGetLastError();
        return fail(DRIFT_PT_ENODEV, "no HIP device visible (GPU engine not built for this host)");
    }
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(DRIFT_PT_ENODEV, "cannot query the HIP device");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(DRIFT_PT_ENODEV, std::string("device is ") + prop.gcnArchName + ", engine built for gfx950");
    return 0;
}

// Bound of one grid-barrier wait of the persistent engines, in device wall-clock ticks: 20 s, or
// DRIFT_PT_BARRIER_TIMEOUT_MS milliseconds when that is set (0 makes the first wait that has to
// spin give up, so the abort path can be exercised from the host; every wave still exits).
unsigned long long barrier_timeout_ticks() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) {
        (void)hipGetLastError();
        khz = 100000;
    }
    unsigned long long ms = 20000ull;
    if (const char* e = std::getenv("DRIFT_PT_BARRIER_TIMEOUT_MS")) {
        char* end = nullptr;
        const unsigned long long v = std::strtoull(e, &end, 10);
        if (end != e && *end == 0) ms = std::min(v, 3600000ull);
    }
    return (unsigned long long)khz * ms;
}

// Device copies of one instance's arrays, and the PtInstance pointing at them.
// Device copies of many instances' arrays in one allocation: every input array is packed into a
// host staging blob at 256-byte aligned offsets and copied once; outputs get their own aligned
// regions in the same allocation (statistics blocks cleared by one memset).
struct Packer {
    std::vector<uint8_t> blob;
    size_t add(const void* src, size_t bytes) {
        const size_t off = (blob.size() + 255) & ~(size_t)255;
        blob.resize(off + std::max<size_t>(bytes, 1), 0);
        if (bytes) std::memcpy(blob.data() + off, src, bytes);
        return off;
    }
    template <class T>
    size_t add(const std::vector<T>& v) { return add(v.data(), v.size() * sizeof(T)); }
};

// First-chunk records (PtInstance::ell): G records per position, row-major, zero-padded,
// positions padded to a multiple of 4, the neighbour stored as position * mult: its LDS spin offset
// on the LDS engine (mult = Rp; below 2^24 since the spins fit one workgroup's LDS), its spin word
// on the word-per-position CSR engine (mult = 1; int8 records need n < 2^24, as the CSR does).
// Each colour class's quads (a quad is 4 consecutive positions, Philox-keyed by its index; the
// quads of class [cs, ce) are cs/4 .. (ce-1)/4, the boundary quads in both their classes) ordered
// by the longest of their rows in the class, longest first (ties by index), the classes in order.
// Any thread may take any quad of a class, so the order changes who sweeps a quad, not what it
// draws or decides.
std::vector<int32_t> build_qorder(const HostInst& H) {
    std::vector<int32_t> out;
    out.reserve((size_t)H.n / 4 + (size_t)H.k + 1);
    for (int c = 0; c < H.k; ++c) {
        const int cs = H.color_ptr[c], ce = H.color_ptr[c + 1];
        if (ce <= cs) continue;
        const int q0 = cs >> 2, q1 = (ce - 1) >> 2;
        std::vector<std::pair<int, int32_t>> key;
        key.reserve((size_t)(q1 - q0 + 1));
        for (int q = q0; q <= q1; ++q) {
            int m = 0;
            for (int p = std::max(4 * q, cs); p < std::min(4 * q + 4, ce); ++p)
                m = std::max(m, H.row_ptr[p + 1] - H.row_ptr[p]);
            key.push_back({-m, q});
        }
        std::sort(key.begin(), key.end());
        for (const auto& kv : key) out.push_back(kv.second);
    }
    return out;
}

std::vector<uint32_t> build_ell(const HostInst& H, int G, uint32_t mult) {
    const bool i8 = H.wm == PT_WM_I8;
    const int ew = i8 ? 1 : 2;
    const size_t n4 = ((size_t)H.n + 3) & ~(size_t)3;
    std::vector<uint32_t> ell(n4 * G * ew, 0u);
    for (int p = 0; p < H.n; ++p) {
        const int t0 = H.row_ptr[p], deg = H.row_ptr[p + 1] - t0;
        for (int u = 0; u < std::min(deg, G); ++u) {
            uint32_t* dst = &ell[((size_t)p * G + u) * ew];
            const uint32_t* src = &H.edge[(size_t)(t0 + u) * ew];
            if (i8) {
                const uint32_t off = (src[0] >> 8) * mult;
                dst[0] = (off << 8) | (src[0] & 0xffu);
            } else {
                dst[0] = src[0] * mult;
                dst[1] = src[1];
            }
        }
    }
    return ell;
}

// Field values zero-padded to a multiple of 4 positions (the LDS engine reads a quad's four at
// once).
template <class T>
std::vector<T> pad4(const std::vector<T>& v) {
    std::vector<T> out((v.size() + 3) & ~(size_t)3, (T)0);
    std::copy(v.begin(), v.end(), out.begin());
    return out;
}

// rgs (split engine): slots per workgroup of each instance; the first-chunk records then hold
// neighbour offsets for groups of that many slots, and img holds one image per group, and each
// instance gets its energy mailboxes (all bits set).
int split_thr_words(const HostInst& H, int rg);

// cl (split engine): also each instance's CSR as 16-bit records (position << 4 | weight & 15).
// nps (split engine, a launch of the kernel with parts): parts per group of each instance; every
// instance gets its spin words and flags (zero) and part mailboxes (all bits set), those of one
// part too, whose energies go through the part mailboxes in such a launch.
int upload_many(DevArena& A, const std::vector<HostInst>& hs, std::vector<PtInstance>& insts,
                bool force_final, bool with_ell, const std::vector<int>* rgs = nullptr,
                bool cl = false, const std::vector<int>* nps = nullptr) {
    const size_t m = hs.size();
    auto rg_of = [&](size_t i) { return rgs ? (*rgs)[i] : hs[i].Rp; };
    auto ng_of = [&](size_t i) { return rgs ? (hs[i].R + (*rgs)[i] - 1) / (*rgs)[i] : 1; };
    auto np_of = [&](size_t i) { return nps ? (*nps)[i] : 1; };
    auto nw_of = [&](size_t i) { return (hs[i].n + 31) / 32; };
    struct Off { size_t rp, edge, ell, cptr, orig, h, beta, thr, jb, crec, qord; };
    std::vector<Off> off(m);
    Packer in;
    for (size_t i = 0; i < m; ++i) {
        const HostInst& H = hs[i];
        Off& o = off[i];
        o.rp = in.add(H.row_ptr);
        o.edge = in.add(H.edge);
        o.ell = with_ell ? in.add(build_ell(H, H.gq, (uint32_t)rg_of(i))) : SIZE_MAX;
        o.qord = with_ell && H.wm == PT_WM_I8 && H.maxdeg > H.gq ? in.add(build_qorder(H))
                                                                   : SIZE_MAX;
        o.cptr = in.add(H.color_ptr);
        o.orig = in.add(H.orig);
        o.h = !H.has_h ? SIZE_MAX
              : with_ell ? (H.wm == PT_WM_F32 ? in.add(pad4(H.hf)) : in.add(pad4(H.h32)))
                         : (H.wm == PT_WM_F32 ? in.add(H.hf) : in.add(H.h32));
        o.beta = in.add(H.beta);
        o.thr = H.wm != PT_WM_F32 ? in.add(H.thr) : SIZE_MAX;
        o.jb = H.ldim ? in.add(H.jbits) : SIZE_MAX;
        o.crec = SIZE_MAX;
        if (cl) {
            std::vector<uint16_t> rec(H.edge.size());
            for (size_t t = 0; t < rec.size(); ++t)
                rec[t] = (uint16_t)(((H.edge[t] >> 8) << 4) | (H.edge[t] & 0xfu));
            o.crec = in.add(rec);
        }
    }
    // outputs after the inputs
    struct OutOff { size_t best, fin, hist, trace, out, img, e0, box, pbox, pbits, pflag; };
    std::vector<OutOff> oo(m);
    size_t cursor = (in.blob.size() + 255) & ~(size_t)255;
    const size_t out_begin = cursor;
    auto take = [&](size_t bytes) {
        const size_t at = cursor;
        cursor = (cursor + std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        return at;
    };
    for (size_t i = 0; i < m; ++i) {
        const HostInst& H = hs[i];
        oo[i].out = take(sizeof(int64_t) * (size_t)pt_out_len(H.R));
    }
    const size_t out_end = cursor;   // statistics blocks are the region that must start at zero
    for (size_t i = 0; i < m; ++i) {
        oo[i].box = rgs ? take(sizeof(uint64_t) * (4 * (size_t)hs[i].Rp + 2)) : SIZE_MAX;
        oo[i].pbox = nps ? take(sizeof(uint64_t) * 8 * (size_t)np_of(i) * hs[i].Rp) : SIZE_MAX;
    }
    const size_t box_end = cursor;   // mailboxes start with every bit set
    for (size_t i = 0; i < m; ++i) {
        oo[i].pbits = nps && np_of(i) > 1
                          ? take(sizeof(uint64_t) * 2 * (size_t)hs[i].Rp * nw_of(i)) : SIZE_MAX;
        oo[i].pflag = nps && np_of(i) > 1
                          ? take(sizeof(uint64_t) * PT_PFLAG_STRIDE * (size_t)ng_of(i) * np_of(i))
                          : SIZE_MAX;
    }
    const size_t bits_end = cursor;  // part spin words start at step 0
    for (size_t i = 0; i < m; ++i) {
        const HostInst& H = hs[i];
        oo[i].best = take((size_t)H.n);
        const bool fin = force_final || (H.flags & DRIFT_PT_RECORD_FINAL);
        oo[i].fin = fin ? take((size_t)H.n * H.R) : SIZE_MAX;
        oo[i].hist = ((H.flags & DRIFT_PT_RECORD_HISTORY) && H.n_rounds > 0)
                         ? take(sizeof(double) * (size_t)H.n_rounds) : SIZE_MAX;
        oo[i].trace = ((H.flags & DRIFT_PT_RECORD_TRACE) && H.n_rounds > 0)
                          ? take(sizeof(double) * (size_t)H.n_rounds * H.R) : SIZE_MAX;
        oo[i].img = with_ell ? take(pt_lds_img_bytes(H.n, rg_of(i)) * (size_t)ng_of(i)) : SIZE_MAX;
        oo[i].e0 = with_ell ? take(sizeof(int64_t) * (size_t)H.Rp) : SIZE_MAX;
    }
    uint8_t* base;
    PT_HIP(A.alloc(&base, cursor));
    PT_HIP(hipMemcpy(base, in.blob.data(), in.blob.size(), hipMemcpyHostToDevice));
    PT_HIP(hipMemset(base + out_begin, 0, out_end - out_begin));
    if (box_end > out_end) PT_HIP(hipMemset(base + out_end, 0xFF, box_end - out_end));
    if (bits_end > box_end) PT_HIP(hipMemset(base + box_end, 0, bits_end - box_end));
    insts.assign(m, PtInstance{});
    for (size_t i = 0; i < m; ++i) {
        const HostInst& H = hs[i];
        const Off& o = off[i];
        PtInstance& I = insts[i];
        I.row_ptr = (const int32_t*)(base + o.rp);
        I.edge = base + o.edge;
        I.ell = o.ell == SIZE_MAX ? nullptr : base + o.ell;
        I.qorder = o.qord == SIZE_MAX ? nullptr : (const int32_t*)(base + o.qord);
        I.color_ptr = (const int32_t*)(base + o.cptr);
        I.orig = (const int32_t*)(base + o.orig);
        I.h = o.h == SIZE_MAX ? nullptr : base + o.h;
        I.beta = (const double*)(base + o.beta);
        I.thr = o.thr == SIZE_MAX ? nullptr : (const uint32_t*)(base + o.thr);
        I.jbits = o.jb == SIZE_MAX ? nullptr : (const uint8_t*)(base + o.jb);
        I.best_s = (int8_t*)(base + oo[i].best);
        I.final_s = oo[i].fin == SIZE_MAX ? nullptr : (int8_t*)(base + oo[i].fin);
        I.hist = oo[i].hist == SIZE_MAX ? nullptr : (double*)(base + oo[i].hist);
        I.trace = oo[i].trace == SIZE_MAX ? nullptr : (double*)(base + oo[i].trace);
        I.out = (int64_t*)(base + oo[i].out);
        I.img = oo[i].img == SIZE_MAX ? nullptr : (int8_t*)(base + oo[i].img);
        I.e0 = oo[i].e0 == SIZE_MAX ? nullptr : (int64_t*)(base + oo[i].e0);
        I.unit = H.unit;
        I.k0 = (uint32_t)H.seed;
        I.k1 = (uint32_t)(H.seed >> 32);
        I.n = H.n; I.k = H.k; I.R = H.R; I.Rp = H.Rp; I.n_rounds = H.n_rounds; I.S = H.S;
        I.TQ = H.TQ; I.wm = H.wm;
        I.ldim = H.ldim; I.lL = H.lL;
        I.gq = H.gq;
        I.ell_rest = H.maxdeg > H.gq ? 1 : 0;
        I.has_target = H.has_target;
        I.target_units = H.target_units;
        I.box = oo[i].box == SIZE_MAX ? nullptr : (uint64_t*)(base + oo[i].box);
        I.rg = rg_of(i);
        I.ng = ng_of(i);
        I.crec = o.crec == SIZE_MAX ? nullptr : (const uint16_t*)(base + o.crec);
        I.nnz = H.nnz;
        I.thr_lds = rgs && split_thr_words(H, rg_of(i)) > 0 ? 1 : 0;
        I.pbox = oo[i].pbox == SIZE_MAX ? nullptr : (uint64_t*)(base + oo[i].pbox);
        I.pbits = oo[i].pbits == SIZE_MAX ? nullptr : (uint64_t*)(base + oo[i].pbits);
        I.pflag = oo[i].pflag == SIZE_MAX ? nullptr : (uint64_t*)(base + oo[i].pflag);
        I.np = np_of(i);
        I.nw = nw_of(i);
    }
    return 0;
}

int upload_instance(DevArena& A, const HostInst& H, PtInstance& I, bool want_final, bool want_hist) {
    (void)want_hist;   // history follows H.flags
    std::vector<HostInst> one{H};
    std::vector<PtInstance> out;
    int rc = upload_many(A, one, out, want_final, false);
    if (rc) return rc;
    I = out[0];
    return 0;
}

void clear_result(drift_pt_result* r) {
    r->best_E = 0.0; r->best_E_units = 0; r->best_round = 0; r->best_rung = 0; r->best_slot = 0;
    r->engine_used = 0; r->seconds = 0.0; r->attempts = 0.0;
    r->kernel_seconds = 0.0; r->kernel_attempts = 0.0; r->kernel_launches = 0;
    r->other_kernel_seconds = 0.0; r->other_kernel_launches = 0; r->rounds_run = 0;
    r->block_threads_used = 0; r->slots_per_group_used = 0; r->couplings_in_lds_used = 0;
    r->parts_per_group_used = 0;
}

// Copy results of one instance back and undo the colour relabelling.
// mirror (optional): a host copy of device memory starting at device address mbase that covers
// this instance's outputs; reads come from it instead of one device copy each.
int fetch_results(const HostInst& H, const PtInstance& I, const std::vector<int64_t>& out,
                  const std::vector<int8_t>& final_pos_major, drift_pt_result* res,
                  const uint8_t* mirror = nullptr, const uint8_t* mbase = nullptr) {
    const int n = H.n, R = H.R;
    auto get = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (!mirror) return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
        std::memcpy(dst, mirror + ((const uint8_t*)src - mbase), bytes);
        return hipSuccess;
    };
    std::vector<int8_t> best((size_t)n);
    PT_HIP(get(best.data(), I.best_s, (size_t)n));
    for (int p = 0; p < n; ++p) res->best_s[H.orig[p]] = best[p];
    res->best_E_units = out[0];
    res->best_E = (double)out[0] * H.unit;
    res->best_round = (int32_t)out[1];
    res->best_rung = (int32_t)out[2];
    res->best_slot = (int32_t)out[3];
    const int ran = (int)std::max<int64_t>(0, std::min<int64_t>(out[5], H.n_rounds));
    res->rounds_run = ran;
    int64_t* o = const_cast<int64_t*>(out.data());
    for (int r = 0; r < R; ++r) {
        if (r + 1 < R) {
            if (res->swap_tried) res->swap_tried[r] = pt_out_arr(o, R, PT_OUT_TRIED)[r];
            if (res->swap_accepted) res->swap_accepted[r] = pt_out_arr(o, R, PT_OUT_ACC)[r];
        }
        if (res->round_trips) res->round_trips[r] = (int32_t)pt_out_arr(o, R, PT_OUT_TRIPS)[r];
        if (res->rung_up) res->rung_up[r] = pt_out_arr(o, R, PT_OUT_UP)[r];
        if (res->rung_down) res->rung_down[r] = pt_out_arr(o, R, PT_OUT_DOWN)[r];
        if (res->rung_of) res->rung_of[r] = (int32_t)pt_out_arr(o, R, PT_OUT_RUNG)[r];
        if (res->final_E_units) res->final_E_units[r] = pt_out_arr(o, R, PT_OUT_E)[r];
    }
    if (res->final_spins && (H.flags & DRIFT_PT_RECORD_FINAL)) {
        if (!final_pos_major.empty()) {
            for (int a = 0; a < R; ++a)
                for (int p = 0; p < n; ++p)
                    res->final_spins[(size_t)a * n + H.orig[p]] = final_pos_major[(size_t)p * H.Rp + a];
        } else {
            std::vector<int8_t> fs((size_t)n * R);
            PT_HIP(get(fs.data(), I.final_s, fs.size()));
            for (int a = 0; a < R; ++a)
                for (int p = 0; p < n; ++p)
                    res->final_spins[(size_t)a * n + H.orig[p]] = fs[(size_t)a * n + p];
        }
    }
    // rounds after a reached target were not run: the history keeps the best, the trace is NaN
    if (res->best_history && (H.flags & DRIFT_PT_RECORD_HISTORY) && H.n_rounds > 0) {
        if (ran > 0) PT_HIP(get(res->best_history, I.hist, sizeof(double) * (size_t)ran));
        for (int r = ran; r < H.n_rounds; ++r) res->best_history[r] = res->best_E;
    }
    if (res->rung_energy_trace && (H.flags & DRIFT_PT_RECORD_TRACE) && H.n_rounds > 0) {
        if (ran > 0) PT_HIP(get(res->rung_energy_trace, I.trace, sizeof(double) * (size_t)ran * R));
        for (size_t x = (size_t)ran * R; x < (size_t)H.n_rounds * R; ++x)
            res->rung_energy_trace[x] = std::nan("");
    }
    res->attempts = (double)R * (double)ran * (double)H.S * (double)n;
    return 0;
}

int choose_lds_block(int Rp, int requested) {
    int b = requested > 0 ? requested : 256;
    b = std::max(b, Rp);
    b = ((b + Rp - 1) / Rp) * Rp;
    b = ((b + 63) / 64) * 64;
    if (b % Rp) b = Rp;   // Rp > 64 is a multiple of 64 already
    return std::min(b, 1024);
}

// fixed (optional): carve every device buffer from this arena's fixed region and run on the null
// stream, timed on the host, so the call creates no device memory, stream or event.
int run_lds(const std::vector<HostInst>& hs, const std::vector<drift_pt_result*>& results,
            int block_req, DevArena* fixed = nullptr) {
    DevArena own;
    DevArena& A = fixed ? *fixed : own;
    const int n_inst = (int)hs.size();
    std::vector<PtInstance> insts((size_t)n_inst);
    size_t lds = 0;
    int maxRp = 1;
    int rc = upload_many(A, hs, insts, false, true);
    if (rc) return rc;
    for (int i = 0; i < n_inst; ++i) {
        lds = std::max(lds, pt_lds_bytes(hs[i].n, hs[i].Rp, hs[i].k));
        maxRp = std::max(maxRp, hs[i].Rp);
    }
    if (lds + 1024 > (size_t)device_lds_limit())
        return fail(DRIFT_PT_EINVAL, "instance does not fit one workgroup's LDS");
    // largest work first (longest-processing-time order for the persistent queue)
    std::vector<int32_t> order((size_t)n_inst);
    std::iota(order.begin(), order.end(), 0);
    auto work = [&](int i) {
        const HostInst& H = hs[i];
        return (double)H.R * H.n_rounds * std::max(H.S, 1) * (H.n + H.nnz);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    PtInstance* d_insts;
    int32_t *d_order, *d_queue;
    PT_HIP(A.upload(&d_insts, insts));
    PT_HIP(A.upload(&d_order, order));
    PT_HIP(A.alloc(&d_queue, 2));   // work queues of the init launch and of the timed launch
    // A few instances leave most of the chip idle, so their workgroups are sized to sweep the
    // largest colour class in one pass (each pass is a dependent chain of loads); batches that
    // fill the chip keep 256 threads.
    int req = block_req;
    if (req <= 0 && n_inst < 64) {
        int quads = 1;
        for (const HostInst& H : hs)
            for (int c = 0; c < H.k; ++c)
                quads = std::max(quads, (H.color_ptr[c + 1] - H.color_ptr[c] + 3) / 4 + 1);
        req = std::max(256, std::min(1024, quads * maxRp));
    }
    const int block = choose_lds_block(maxRp, req);
    for (const HostInst& H : hs)
        if (block % H.Rp) return fail(DRIFT_PT_EINVAL, "replica counts in one batch need a common block size");
    const int grid = std::min(n_inst, 256 * 8);
    StreamGuard sg;
    float ms = 0.0f;
    // The initial spins and energies are drawn by a launch of their own, so the timed launch is
    // the round loop alone (cuda/ising_pt.cu:213-222 times the rounds, not init_kernel).
    const int wm = hs[0].wm;
    if (fixed) {
        PT_HIP(hipMemsetAsync(d_queue, 0, 2 * sizeof(int32_t), 0));
        PT_HIP(pt_launch_lds(d_insts, d_order, n_inst, d_queue, wm, block, grid, lds, true, 0));
        PT_HIP(hipStreamSynchronize(0));
        const auto t0 = std::chrono::steady_clock::now();
        PT_HIP(pt_launch_lds(d_insts, d_order, n_inst, d_queue + 1, wm, block, grid, lds, false, 0));
        PT_HIP(hipStreamSynchronize(0));
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                 .count();
    } else {
        PT_HIP(hipStreamCreate(&sg.s));
        PT_HIP(hipEventCreate(&sg.e0));
        PT_HIP(hipEventCreate(&sg.e1));
        PT_HIP(hipMemsetAsync(d_queue, 0, 2 * sizeof(int32_t), sg.s));
        PT_HIP(pt_launch_lds(d_insts, d_order, n_inst, d_queue, wm, block, grid, lds, true, sg.s));
        PT_HIP(hipEventRecord(sg.e0, sg.s));
        PT_HIP(pt_launch_lds(d_insts, d_order, n_inst, d_queue + 1, wm, block, grid, lds, false, sg.s));
        PT_HIP(hipEventRecord(sg.e1, sg.s));
        PT_HIP(hipStreamSynchronize(sg.s));
        PT_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));
    }
    // Every instance's outputs live in the one arena allocation: one copy brings them all back.
    const uint8_t *lo = nullptr, *hi = nullptr;
    auto span = [&](const void* p, size_t bytes) {
        if (!p || !bytes) return;
        const uint8_t* a = (const uint8_t*)p;
        if (!lo || a < lo) lo = a;
        if (!hi || a + bytes > hi) hi = a + bytes;
    };
    for (int i = 0; i < n_inst; ++i) {
        const HostInst& H = hs[i];
        const PtInstance& I = insts[i];
        span(I.out, sizeof(int64_t) * (size_t)pt_out_len(H.R));
        span(I.best_s, (size_t)H.n);
        if (H.flags & DRIFT_PT_RECORD_FINAL) span(I.final_s, (size_t)H.n * H.R);
        if (H.flags & DRIFT_PT_RECORD_HISTORY) span(I.hist, sizeof(double) * (size_t)H.n_rounds);
        if (H.flags & DRIFT_PT_RECORD_TRACE)
            span(I.trace, sizeof(double) * (size_t)H.n_rounds * H.R);
    }
    std::vector<uint8_t> mirror(lo ? (size_t)(hi - lo) : 0);
    if (lo) PT_HIP(hipMemcpy(mirror.data(), lo, mirror.size(), hipMemcpyDeviceToHost));
    double total = 0.0;
    for (const HostInst& Hx : hs) total += (double)Hx.R * Hx.n_rounds * Hx.S * Hx.n;
    for (int i = 0; i < n_inst; ++i) {
        const HostInst& H = hs[i];
        std::vector<int64_t> out((size_t)pt_out_len(H.R));
        std::memcpy(out.data(), mirror.data() + ((const uint8_t*)insts[i].out - lo),
                    out.size() * sizeof(int64_t));
        int rc = fetch_results(H, insts[i], out, {}, results[i], mirror.data(), lo);
        if (rc) return rc;
        results[i]->seconds = ms * 1e-3;
        results[i]->engine_used = DRIFT_PT_ENGINE_LDS;
        results[i]->block_threads_used = block;
        results[i]->slots_per_group_used = H.Rp;
        results[i]->kernel_seconds = ms * 1e-3;      // one persistent launch does everything
        results[i]->kernel_launches = 1;
        results[i]->kernel_attempts = total;
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// split engine: an instance's replica slots in groups, one workgroup's LDS each
// ------------------------------------------------------------------------------------------
int device_cus() {
    static int cached = -1;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (cached < 0) {
        hipDeviceProp_t prop;
        int dev = 0;
        cached = 256;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cached = std::max(1, prop.multiProcessorCount);
    }
    return cached;
}

// Largest power-of-two number of slots per workgroup (at most 64 and Rp) whose spins and ladder
// arrays fit one workgroup's LDS; 0 when not even one slot fits.
// Threshold words a split workgroup of rg slots stages in LDS each round (integer couplings,
// when they fit 16 KB), or 0.
int split_thr_words(const HostInst& H, int rg) {
    return H.wm != PT_WM_F32 && (int64_t)rg * H.TQ * 4 <= 16384 ? rg * H.TQ : 0;
}

size_t split_lds(const HostInst& H, int rg, bool cl, bool parts = false) {
    return pt_split_lds_bytes(H.n, rg, H.Rp, H.k, split_thr_words(H, rg), cl, H.nnz, H.has_h,
                              parts);
}

// 32-position words of the largest colour class.  Parts per group are at most this, so every part
// owns a word of that class, which the split kernel's spin-word buffer reuse relies on.
int class_words_max(const HostInst& H) {
    int best = 1;
    for (int c = 0; c < H.k; ++c) {
        const int cs = H.color_ptr[c], ce = H.color_ptr[c + 1];
        if (ce > cs) best = std::max(best, ((ce - 1) >> 5) + 1 - (cs >> 5));
    }
    return best;
}

// Parts per group of each instance: the requested count where one is given, else 1.
std::vector<int> split_nps(const std::vector<HostInst>& hs) {
    std::vector<int> out(hs.size(), 1);
    for (size_t i = 0; i < hs.size(); ++i)
        if (hs[i].np_req > 0) out[i] = hs[i].np_req;
    return out;
}

int split_rg_max(const HostInst& H, bool cl = false) {
    if (cl && !H.cl_ok) return 0;
    for (int rg = std::min(64, H.Rp); rg >= 1; rg >>= 1)
        if (split_lds(H, rg, cl) + 1024 <= (size_t)device_lds_limit()) return rg;
    return 0;
}

bool split_fits(const HostInst& H) { return split_rg_max(H) >= 1; }

// Couplings in LDS for a launch: asked for (DRIFT_PT_COUPLINGS_LDS) by every instance.
bool wants_cl(const std::vector<HostInst>& hs) {
    for (const HostInst& H : hs)
        if (!(H.flags & DRIFT_PT_COUPLINGS_LDS)) return false;
    return !hs.empty();
}

// Fewest slots per group (a power of two) that keep an instance's groups within one per compute
// unit: every group of an instance is resident at once, and one workgroup per compute unit
// always is.
int split_rg_min(const HostInst& H) {
    int rg = 1;
    while (rg < 64 && (H.R + rg - 1) / rg > device_cus()) rg <<= 1;
    return rg;
}

// Slots per workgroup for a set of instances run together: the requested size where one is
// given, else the largest common cap c (a power of two) under which the groups number at least
// two per compute unit, each instance taking min(max(c, split_rg_min), its largest fitting size);
// c = 1 when no cap reaches that.  Larger groups share each position's edge records among more
// lanes; more groups keep more compute units busy.
std::vector<int> split_rgs(const std::vector<HostInst>& hs, bool cl = false) {
    std::vector<int> out(hs.size(), 1);
    const int64_t want = 2 * (int64_t)device_cus();
    auto pick = [&](const HostInst& H, int c) {
        return H.rg_req > 0 ? H.rg_req
                            : std::max(1, std::min(std::max(c, split_rg_min(H)), split_rg_max(H, cl)));
    };
    int cap = 1;
    for (int c = 64; c >= 1; c >>= 1) {
        int64_t groups = 0;
        for (const HostInst& H : hs) {
            const int rg = pick(H, c);
            groups += (H.R + rg - 1) / rg;
        }
        if (groups >= want) {
            cap = c;
            break;
        }
    }
    for (size_t i = 0; i < hs.size(); ++i) out[i] = pick(hs[i], cap);
    return out;
}

// Workgroup size of the split engine: the requested one, else by the LDS a workgroup takes, so
// the workgroups a compute unit holds fill its 32 wave slots (1024 threads when at most two fit,
// down to 256 when eight or more do).  On 64 3-regular instances of N = 65536 with R = 32,
// 1024-thread workgroups ran at 442 attempts/ns against 351 at 256 threads.
int split_block(size_t lds, int block_req) {
    if (block_req > 0) return std::min(1024, std::max(64, (block_req + 63) / 64 * 64));
    const int per_cu = std::max(1, (int)((size_t)device_lds_limit() / (lds + 1024)));
    int b = 2048 / per_cu;
    b = std::max(256, std::min(1024, b / 64 * 64));
    return b;
}

// Instances on the split engine, all of one weight type: one launch draws every group's initial
// state, then one persistent launch (timed) runs every round of every instance.
int run_split(const std::vector<HostInst>& hs, const std::vector<drift_pt_result*>& results,
              int block_req) {
    DevArena A;
    const int n_inst = (int)hs.size();
    if (n_inst >= (1 << 21)) return fail(DRIFT_PT_EINVAL, "split engine: at most 2^21 instances per batch");
    const bool cl = wants_cl(hs);
    for (const HostInst& H : hs)
        if (cl && !H.cl_ok)
            return fail(DRIFT_PT_EINVAL, "couplings in LDS need int8 couplings with |w| <= 7 and "
                                         "n <= 4096");
    const std::vector<int> rgs = split_rgs(hs, cl);
    const std::vector<int> nps = split_nps(hs);
    const bool parts = std::any_of(nps.begin(), nps.end(), [](int v) { return v > 1; });
    if (parts && cl)
        return fail(DRIFT_PT_EINVAL, "split engine: couplings in LDS do not combine with parts "
                                     "per group");
    size_t lds = 0;
    int max_ng = 1;   // workgroups of the largest instance (groups x parts)
    bool plain = false;
    for (int i = 0; i < n_inst; ++i) {
        const HostInst& H = hs[i];
        const int rg = rgs[i];
        if (rg < 1 || rg > H.Rp || rg > split_rg_max(H, cl))
            return fail(DRIFT_PT_EINVAL, "split engine: " + std::to_string(rg) +
                                             " slots per workgroup do not fit one workgroup's LDS "
                                             "(or exceed the padded replica count)");
        const int ng = (H.R + rg - 1) / rg, np = nps[i];
        if (np > 1) {
            if (np > class_words_max(H))
                return fail(DRIFT_PT_EINVAL, "split engine: " + std::to_string(np) +
                                                 " parts per group exceed the " +
                                                 std::to_string(class_words_max(H)) +
                                                 " 32-position words of the largest colour class");
            if ((int64_t)ng * np > 1024)
                return fail(DRIFT_PT_EINVAL, "split engine: more than 1024 workgroups (groups x "
                                             "parts) for one instance");
            if ((int64_t)H.n_rounds * std::max(H.S, 1) * std::max(H.k, 1) >= ((int64_t)1 << 31))
                return fail(DRIFT_PT_ERANGE, "split engine with parts: more than 2^31 colour steps");
        }
        if (parts && split_lds(H, rg, cl, true) > (size_t)device_lds_limit())
            return fail(DRIFT_PT_EINVAL, "split engine: a group's state with the arrays of parts "
                                         "does not fit one workgroup's LDS");
        lds = std::max(lds, split_lds(H, rg, cl, parts));
        max_ng = std::max(max_ng, ng * np);
        if (H.flags & DRIFT_PT_PLAIN_LAUNCH) plain = true;
        if (hs[i].wm != hs[0].wm) return fail(DRIFT_PT_EINVAL, "split engine: one weight type per launch");
    }
    // groups in parts: 1024 threads unless asked (fewer ran slower at every parts count measured)
    const int block = parts && block_req <= 0 ? 1024 : split_block(lds, block_req);
    const int wm = hs[0].wm;
    int per_cu = 0;
    PT_HIP(pt_split_occupancy(wm, block, lds, &per_cu, cl, parts));
    const int capacity = per_cu * device_cus();
    if (capacity < max_ng)
        return fail(DRIFT_PT_EINVAL, "split engine: the " + std::to_string(max_ng) +
                                         " groups of an instance cannot all be resident (room for " +
                                         std::to_string(capacity) + " workgroups of " +
                                         std::to_string(block) + " threads)");
    std::vector<PtInstance> insts((size_t)n_inst);
    int rc = upload_many(A, hs, insts, false, true, &rgs, cl, parts ? &nps : nullptr);
    if (rc) return rc;
    // Largest work first, the groups of one instance consecutive.  With eight queues (one per
    // XCD: a workgroup takes from queue blockIdx.x mod 8 first) each instance goes, largest
    // first, to the queue with the least work so far, so its groups share an L2.  A workgroup
    // waits only on groups of an instance its queue has started; at most nq (max groups - 1)
    // workgroups can hold such groups at once, so more resident workgroups than that guarantee
    // progress; otherwise one queue.
    std::vector<int32_t> order((size_t)n_inst);
    std::iota(order.begin(), order.end(), 0);
    auto work = [&](int i) {
        const HostInst& H = hs[i];
        return (double)H.n_rounds * std::max(H.S, 1) * (H.n + H.nnz) * H.R;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    // the init launch takes groups, the timed one every (group, part) workgroup
    const int n_work_all = [&] {
        int t = 0;
        for (const PtInstance& I : insts) t += I.ng * I.np;
        return t;
    }();
    const int grid = std::min(n_work_all, capacity);
    const int nq = (n_inst >= 8 && grid > 8 * (max_ng - 1)) ? 8 : 1;
    std::vector<std::vector<int32_t>> qitems((size_t)nq), qinit((size_t)nq);
    std::vector<double> qwork((size_t)nq, 0.0);
    for (int i : order) {
        const int q = (int)(std::min_element(qwork.begin(), qwork.end()) - qwork.begin());
        qwork[q] += work(i);
        for (int g = 0; g < insts[i].ng; ++g) qinit[q].push_back((int32_t)((i << 10) | g));
        for (int w = 0; w < insts[i].ng * insts[i].np; ++w)
            qitems[q].push_back((int32_t)((i << 10) | w));
    }
    std::vector<int32_t> items, qoff(1, 0), items0, qoff0(1, 0);
    for (int q = 0; q < nq; ++q) {
        items.insert(items.end(), qitems[q].begin(), qitems[q].end());
        qoff.push_back((int32_t)items.size());
        items0.insert(items0.end(), qinit[q].begin(), qinit[q].end());
        qoff0.push_back((int32_t)items0.size());
    }
    const int n_items = (int)items.size(), n_items0 = (int)items0.size();
    PtInstance* d_insts;
    int32_t *d_items, *d_items0, *d_queue, *d_abort, *d_qoff, *d_qoff0;
    PT_HIP(A.upload(&d_insts, insts));
    PT_HIP(A.upload(&d_items, items));
    PT_HIP(A.upload(&d_qoff, qoff));
    PT_HIP(A.upload(&d_items0, items0));
    PT_HIP(A.upload(&d_qoff0, qoff0));
    PT_HIP(A.alloc(&d_queue, 2 * (size_t)nq));   // the init launch's counters, then the timed one's
    PT_HIP(A.alloc(&d_abort, 1));
    SplitArgs SA{d_abort, barrier_timeout_ticks(), d_qoff, nq};
    const SplitArgs SA0{d_abort, barrier_timeout_ticks(), d_qoff0, nq};
    StreamGuard sg;
    PT_HIP(hipStreamCreate(&sg.s));
    PT_HIP(hipEventCreate(&sg.e0));
    PT_HIP(hipEventCreate(&sg.e1));
    PT_HIP(hipMemsetAsync(d_queue, 0, 2 * (size_t)nq * sizeof(int32_t), sg.s));
    PT_HIP(hipMemsetAsync(d_abort, 0, sizeof(int32_t), sg.s));
    PT_HIP(pt_launch_split(d_insts, d_items0, n_items0, d_queue, wm, block, grid, lds, true, SA0,
                           false, sg.s, cl));
    PT_HIP(hipEventRecord(sg.e0, sg.s));
    PT_HIP(pt_launch_split(d_insts, d_items, n_items, d_queue + nq, wm, block, grid, lds, false, SA,
                           !plain, sg.s, cl, parts));
    PT_HIP(hipEventRecord(sg.e1, sg.s));
    PT_HIP(hipStreamSynchronize(sg.s));
    float ms = 0.0f;
    PT_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));
    int32_t aborted = 0;
    PT_HIP(hipMemcpy(&aborted, d_abort, sizeof(aborted), hipMemcpyDeviceToHost));
    if (aborted) return fail(DRIFT_PT_EHIP, "split engine: a mailbox wait exceeded its bound");
    // every instance's outputs in one copy, as on the LDS engine
    const uint8_t *lo = nullptr, *hi = nullptr;
    auto span = [&](const void* p, size_t bytes) {
        if (!p || !bytes) return;
        const uint8_t* a = (const uint8_t*)p;
        if (!lo || a < lo) lo = a;
        if (!hi || a + bytes > hi) hi = a + bytes;
    };
    for (int i = 0; i < n_inst; ++i) {
        const HostInst& H = hs[i];
        const PtInstance& I = insts[i];
        span(I.out, sizeof(int64_t) * (size_t)pt_out_len(H.R));
        span(I.best_s, (size_t)H.n);
        if (H.flags & DRIFT_PT_RECORD_FINAL) span(I.final_s, (size_t)H.n * H.R);
        if (H.flags & DRIFT_PT_RECORD_HISTORY) span(I.hist, sizeof(double) * (size_t)H.n_rounds);
        if (H.flags & DRIFT_PT_RECORD_TRACE)
            span(I.trace, sizeof(double) * (size_t)H.n_rounds * H.R);
    }
    std::vector<uint8_t> mirror(lo ? (size_t)(hi - lo) : 0);
    if (lo) PT_HIP(hipMemcpy(mirror.data(), lo, mirror.size(), hipMemcpyDeviceToHost));
    double total = 0.0;
    for (const HostInst& Hx : hs) total += (double)Hx.R * Hx.n_rounds * Hx.S * Hx.n;
    for (int i = 0; i < n_inst; ++i) {
        const HostInst& H = hs[i];
        std::vector<int64_t> out((size_t)pt_out_len(H.R));
        std::memcpy(out.data(), mirror.data() + ((const uint8_t*)insts[i].out - lo),
                    out.size() * sizeof(int64_t));
        rc = fetch_results(H, insts[i], out, {}, results[i], mirror.data(), lo);
        if (rc) return rc;
        results[i]->seconds = ms * 1e-3;
        results[i]->engine_used = DRIFT_PT_ENGINE_SPLIT;
        results[i]->block_threads_used = block;
        results[i]->slots_per_group_used = rgs[i];
        results[i]->couplings_in_lds_used = cl ? 1 : 0;
        results[i]->parts_per_group_used = nps[i];
        results[i]->kernel_seconds = ms * 1e-3;
        results[i]->kernel_launches = 1;
        results[i]->kernel_attempts = total;
    }
    return 0;
}

// Results of the engines that keep their state in a GState: statistics block, best record, final
// energies and rungs from the buffers of the parity of the rounds run.
int finish_global(const HostInst& H, const PtInstance& I, const GState& G, int rounds_run,
                  drift_pt_result* res) {
    const int Rp = H.Rp;
    const int fin = rounds_run & 1;
    std::vector<int64_t> out((size_t)pt_out_len(H.R));
    PT_HIP(hipMemcpy(out.data(), I.out, out.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t best[5];
    PT_HIP(hipMemcpy(best, G.best + 5 * fin, sizeof(best), hipMemcpyDeviceToHost));
    std::vector<int64_t> E((size_t)Rp);
    std::vector<int32_t> rung((size_t)Rp);
    PT_HIP(hipMemcpy(E.data(), G.Eb + (size_t)fin * Rp, sizeof(int64_t) * Rp, hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(rung.data(), G.rung + (size_t)fin * Rp, sizeof(int32_t) * Rp, hipMemcpyDeviceToHost));
    for (int x = 0; x < 5; ++x) out[x] = best[x];
    out[5] = rounds_run;
    for (int r = 0; r < H.R; ++r) {
        pt_out_arr(out.data(), H.R, PT_OUT_E)[r] = E[r];
        pt_out_arr(out.data(), H.R, PT_OUT_RUNG)[r] = rung[r];
    }
    std::vector<int8_t> fs;
    if (H.flags & DRIFT_PT_RECORD_FINAL) {
        fs.resize((size_t)H.n * Rp);
        PT_HIP(hipMemcpy(fs.data(), G.sp, fs.size(), hipMemcpyDeviceToHost));
    }
    return fetch_results(H, I, out, fs, res);
}

// One run of the per-launch engines (global CSR or lattice), driven step by step so that several
// can share the device: setup() uploads the instance, draws the initial state and captures a
// chunk of rounds into a hipGraph; each step() enqueues one chunk (a graph replay, or direct
// launches of up to a chunk of rounds) on the run's own stream; finish() waits and fetches the
// results.  With a target, every step also copies the device's stop flag into pinned host memory,
// and the flag of the step before is read once that step is done, so the device never waits on
// the host; rounds enqueued after the target was reached return at entry.
struct GlobalRun {
    const HostInst& H;
    drift_pt_result* res;
    int block_req;
    bool lattice;
    DevArena A;
    PtInstance I{};
    GState G{};
    StreamGuard sg;
    GraphGuard gg;
    EventPool pool, other, polls;
    int block = 0, grid = 0, reg = 0, chunk = 0, replays = 0, rest = 0, done_replays = 0,
        done_rest = 0, steps = 0, rounds_run = -1;
    bool bits = false, time_k = false;
    std::vector<int> cgrid;
    int32_t* host_stop = nullptr;   // pinned, one slot per step parity
    std::vector<hipEvent_t> poll_ev;
    int64_t timed = 0, timed_other = 0;
    double timed_attempts = 0.0;

    GlobalRun(const HostInst& h, drift_pt_result* r, int b, bool lat)
        : H(h), res(r), block_req(b), lattice(lat) {}
    ~GlobalRun() {
        if (host_stop) (void)hipHostFree(host_stop);
    }

    int launch_round(int rl, bool timing) {
        for (int j = 0; j < H.S; ++j)
            for (int c = 0; c < H.k; ++c) {
                const bool tk = timing && timed < 4096;
                hipEvent_t a = nullptr, b = nullptr;
                if (tk) {
                    PT_HIP(pool.pair(&a, &b));
                    PT_HIP(hipEventRecord(a, sg.s));
                }
                if (lattice) PT_HIP(pt_launch_g_lattice(I, G, c, j, rl, block, grid, sg.s));
                else PT_HIP(pt_launch_g_color(I, G, H.color_ptr[c], H.color_ptr[c + 1], j, rl,
                                              block, cgrid[c], sg.s));
                if (tk) {
                    PT_HIP(hipEventRecord(b, sg.s));
                    ++timed;
                    timed_attempts += (double)H.R * (H.color_ptr[c + 1] - H.color_ptr[c]);
                }
            }
        const bool tr = timing && timed_other < 4096;
        hipEvent_t a = nullptr, b = nullptr;
        if (tr) {
            PT_HIP(other.pair(&a, &b));
            PT_HIP(hipEventRecord(a, sg.s));
        }
        PT_HIP(pt_launch_g_round_end(I, G, rl, reg, sg.s));
        if (tr) {
            PT_HIP(hipEventRecord(b, sg.s));
            ++timed_other;
        }
        return 0;
    }

    int setup() {
        int rc = upload_instance(A, H, I, false, H.flags & DRIFT_PT_RECORD_HISTORY);
        if (rc) return rc;
        if (!lattice) I.jbits = nullptr;
        const int Rp = H.Rp;
        PT_HIP(A.alloc(&G.sp, (size_t)H.n * Rp));
        PT_HIP(A.alloc(&G.Eb, 2 * (size_t)Rp));
        PT_HIP(A.alloc(&G.dacc, 2 * (size_t)PT_DACC_SPLIT * Rp));
        PT_HIP(A.alloc(&G.rung, 2 * (size_t)Rp));
        PT_HIP(A.alloc(&G.rep, 2 * (size_t)Rp));
        PT_HIP(A.alloc(&G.label, (size_t)Rp));
        PT_HIP(A.alloc(&G.best, 10));
        PT_HIP(A.alloc(&G.round_base, 1));
        if (H.has_target) {
            PT_HIP(A.alloc(&G.stop, 1));
            PT_HIP(hipHostMalloc((void**)&host_stop, 2 * sizeof(int32_t), hipHostMallocDefault));
            host_stop[0] = host_stop[1] = 0;
        }
        PT_HIP(hipStreamCreate(&sg.s));
        PT_HIP(hipEventCreate(&sg.e0));
        PT_HIP(hipEventCreate(&sg.e1));
        PT_HIP(hipMemsetAsync(G.Eb, 0, 2 * sizeof(int64_t) * Rp, sg.s));
        PT_HIP(hipMemsetAsync(G.dacc, 0, 2 * PT_DACC_SPLIT * sizeof(int64_t) * Rp, sg.s));
        PT_HIP(hipMemsetAsync(G.rung, 0, 2 * sizeof(int32_t) * Rp, sg.s));
        PT_HIP(hipMemsetAsync(G.rep, 0, 2 * sizeof(int32_t) * Rp, sg.s));
        PT_HIP(hipMemsetAsync(G.label, 0, sizeof(int32_t) * Rp, sg.s));
        PT_HIP(hipMemsetAsync(G.best, 0, 10 * sizeof(int64_t), sg.s));
        PT_HIP(hipMemsetAsync(G.round_base, 0, sizeof(int32_t), sg.s));
        if (G.stop) PT_HIP(hipMemsetAsync(G.stop, 0, sizeof(int32_t), sg.s));

        block = block_req > 0 ? block_req : 256;
        block = std::max(block, Rp);
        block = std::min(((block + Rp - 1) / Rp) * Rp, 1024);
        int maxc = 0;
        for (int c = 0; c < H.k; ++c) maxc = std::max(maxc, H.color_ptr[c + 1] - H.color_ptr[c]);
        const int64_t threads = (int64_t)((maxc + 3) / 4 + 1) * Rp;
        grid = (int)std::max<int64_t>(1, std::min<int64_t>((threads + block - 1) / block, 2048));
        // CSR colour steps run one attempt per lane: (block / Rp) positions per workgroup;
        // R <= 32 on the CSR engine: spins as one word per position (32 lanes per position)
        bits = !lattice && H.R <= 32 && block % 64 == 0;
        const int ppb = bits ? block / 32 : block / Rp;
        cgrid.assign((size_t)H.k, 1);
        for (int c = 0; c < H.k; ++c)
            cgrid[c] = std::max(1, (H.color_ptr[c + 1] - H.color_ptr[c] + ppb - 1) / ppb);
        reg = std::max(1, std::min(256, (H.n + 1023) / 1024));   // round-end grid
        PT_HIP(pt_launch_g_init(I, G, block, reg, sg.s));   // byte spins, initial energies and state
        if (bits) {
            uint32_t* spw = nullptr;
            PT_HIP(A.alloc(&spw, (size_t)H.n));
            G.spw = spw;
            PT_HIP(pt_launch_g_pack(I, G, false, sg.s));
        }
        if (!lattice) {
            // first-chunk records by position: 4 per row when that covers three rows in four, else 8
            int short_rows = 0;
            for (int p = 0; p < H.n; ++p) short_rows += H.row_ptr[p + 1] - H.row_ptr[p] <= 4;
            I.gq = 4 * (int64_t)short_rows >= 3 * (int64_t)H.n ? 4 : 8;
            I.ell_rest = H.maxdeg > I.gq ? 1 : 0;
            uint32_t* ell = nullptr;
            PT_HIP(A.upload(&ell, build_ell(H, I.gq, 1u)));
            I.ell = ell;
        }
        time_k = (H.flags & DRIFT_PT_TIME_KERNELS) != 0;
        // Chunks of rounds are captured once into a hipGraph (rounds read their index from the
        // device counter, which a tick kernel advances) and replayed; leftovers launch directly.
        // With a target the chunks are short, so little runs past the round that reached it.
        const int per_round = H.S * H.k + 1;
        const int launches = H.has_target ? 256 : 1024;
        chunk = std::max(1, std::min(H.n_rounds, launches / per_round));
        if (!time_k && !(H.flags & DRIFT_PT_NO_GRAPH) && H.n_rounds >= 4) {
            PT_HIP(hipStreamBeginCapture(sg.s, hipStreamCaptureModeThreadLocal));
            for (int rl = 0; rl < chunk; ++rl) {
                const int rc2 = launch_round(rl, false);
                if (rc2) {
                    hipGraph_t junk = nullptr;
                    (void)hipStreamEndCapture(sg.s, &junk);
                    if (junk) (void)hipGraphDestroy(junk);
                    return rc2;
                }
            }
            PT_HIP(pt_launch_g_tick(G, chunk, sg.s));
            PT_HIP(hipStreamEndCapture(sg.s, &gg.graph));
            PT_HIP(hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0));
            replays = H.n_rounds / chunk;
        }
        rest = H.n_rounds - replays * chunk;
        PT_HIP(hipEventRecord(sg.e0, sg.s));
        return 0;
    }

    // The stop flag of step k - 1, once that step is done (with a target; else nothing).
    int poll_previous() {
        if (!H.has_target || steps < 2) return 0;
        const int k = steps - 2;
        PT_HIP(hipEventSynchronize(poll_ev[(size_t)k]));
        const int32_t s = host_stop[k & 1];
        if (s != 0) rounds_run = s - 1;
        return 0;
    }

    // Enqueue the next chunk; *more is false once every round is enqueued or the target was seen.
    int step(bool* more) {
        *more = false;
        if (rounds_run >= 0) return 0;
        if (done_replays < replays) {
            PT_HIP(hipGraphLaunch(gg.exec, sg.s));
            ++done_replays;
        } else if (done_rest < rest) {
            const int n = std::min(chunk, rest - done_rest);
            for (int x = 0; x < n; ++x) {
                const int rc = launch_round(done_rest + x, time_k);
                if (rc) return rc;
            }
            done_rest += n;
        } else {
            return 0;
        }
        if (H.has_target) {
            hipEvent_t e = nullptr;
            PT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            polls.ev.push_back(e);
            poll_ev.push_back(e);
            PT_HIP(hipMemcpyAsync(host_stop + (steps & 1), G.stop, sizeof(int32_t),
                                  hipMemcpyDeviceToHost, sg.s));
            PT_HIP(hipEventRecord(e, sg.s));
        }
        ++steps;
        int rc = poll_previous();
        if (rc) return rc;
        *more = rounds_run < 0 && (done_replays < replays || done_rest < rest);
        return 0;
    }

    int finish() {
        PT_HIP(hipEventRecord(sg.e1, sg.s));
        PT_HIP(hipStreamSynchronize(sg.s));
        if (rounds_run < 0) {
            rounds_run = H.n_rounds;
            if (H.has_target) {
                int32_t s = 0;
                PT_HIP(hipMemcpy(&s, G.stop, sizeof(s), hipMemcpyDeviceToHost));
                if (s != 0) rounds_run = s - 1;
            }
        }
        float ms = 0.0f;
        PT_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));
        double kernel_s = 0.0, other_s = 0.0;
        for (size_t x = 0; x < pool.ev.size(); x += 2) {
            float km = 0.0f;
            PT_HIP(hipEventElapsedTime(&km, pool.ev[x], pool.ev[x + 1]));
            kernel_s += km * 1e-3;
        }
        for (size_t x = 0; x < other.ev.size(); x += 2) {
            float km = 0.0f;
            PT_HIP(hipEventElapsedTime(&km, other.ev[x], other.ev[x + 1]));
            other_s += km * 1e-3;
        }
        if (bits) {   // final spins back to bytes where finish_global reads them
            PT_HIP(pt_launch_g_pack(I, G, true, sg.s));
            PT_HIP(hipStreamSynchronize(sg.s));
        }
        const int rc = finish_global(H, I, G, rounds_run, res);
        if (rc) return rc;
        res->seconds = ms * 1e-3;
        res->engine_used = lattice ? DRIFT_PT_ENGINE_LATTICE : DRIFT_PT_ENGINE_GLOBAL;
        res->block_threads_used = block;
        if (time_k) {
            res->kernel_seconds = kernel_s;
            res->kernel_attempts = timed_attempts;
            res->kernel_launches = timed;
            res->other_kernel_seconds = other_s;
            res->other_kernel_launches = timed_other;
        } else {
            res->kernel_seconds = res->seconds;
            res->kernel_attempts = res->attempts;
            res->kernel_launches = 1;
        }
        return 0;
    }
};

int run_global(const HostInst& H, drift_pt_result* res, int block_req, bool use_lattice) {
    GlobalRun run(H, res, block_req, use_lattice);
    int rc = run.setup();
    if (rc) return rc;
    for (bool more = true; more;) {
        rc = run.step(&more);
        if (rc) return rc;
    }
    return run.finish();
}

// Several per-launch runs sharing the device: up to `width` at a time, each on its own stream, one
// chunk each in turn, so one run's launches fill the gaps between another's.  Every run makes the
// decisions it makes alone.  seconds of each result is its own stream's time; wall_s (optional)
// receives the host time of the whole set.
int run_global_set(const std::vector<const HostInst*>& hs, const std::vector<drift_pt_result*>& rs,
                   int width, double* wall_s) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::unique_ptr<GlobalRun>> active;
    size_t next = 0;
    auto start_next = [&]() -> int {
        const HostInst& H = *hs[next];
        active.emplace_back(new GlobalRun(H, rs[next], 0, H.ldim != 0));
        ++next;
        return active.back()->setup();
    };
    while (next < hs.size() && (int)active.size() < width) {
        const int rc = start_next();
        if (rc) return rc;
    }
    while (!active.empty()) {
        for (size_t i = 0; i < active.size();) {
            bool more = false;
            int rc = active[i]->step(&more);
            if (rc) return rc;
            if (more) {
                ++i;
                continue;
            }
            rc = active[i]->finish();
            if (rc) return rc;
            active.erase(active.begin() + (long)i);
            if (next < hs.size()) {
                rc = start_next();
                if (rc) return rc;
            }
        }
    }
    if (wall_s)
        *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// ------------------------------------------------------------------------------------------
// coop engine: one cooperative launch, each workgroup's share of the rows resident in its LDS
// ------------------------------------------------------------------------------------------
struct CoopPlan {
    int block = 0, grid = 0;
    size_t lds = 0;
    CoopArgs C{};
};

// LDS layout of one coop workgroup when the colour classes are split W ways; returns its bytes.
size_t coop_layout(const HostInst& H, int W, CoopArgs& C) {
    int64_t max_own = 0, max_edges = 0;   // the largest slice over the workgroups
    for (int w = 0; w < W; ++w) {
        int64_t own = 0, edges = 0;
        for (int c = 0; c < H.k; ++c) {
            const int cs = H.color_ptr[c], cz = H.color_ptr[c + 1] - cs;
            const int lo = cs + (int)((int64_t)cz * w / W), hi = cs + (int)((int64_t)cz * (w + 1) / W);
            own += hi - lo;
            edges += H.row_ptr[hi] - H.row_ptr[lo];
        }
        max_own = std::max(max_own, own);
        max_edges = std::max(max_edges, edges);
    }
    const size_t Rp = (size_t)H.Rp, ew = H.wm == PT_WM_I8 ? 1 : 2;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 15) & ~(size_t)15;
        return (uint32_t)at;
    };
    C.off_cptr = take(4 * ((size_t)H.k + 1));
    C.off_lrp = take(4 * ((size_t)max_own + H.k));      // one end offset per colour
    C.off_edge = take(4 * ew * (size_t)max_edges);
    C.off_lsp = take((size_t)max_own * Rp);              // own spins, [own position][slot]
    const size_t thr_bytes = H.wm != PT_WM_F32 ? 4 * (size_t)H.R * H.TQ : 0;
    C.thr_lds = thr_bytes > 0 && thr_bytes <= 32768;
    C.off_thr = take(C.thr_lds ? thr_bytes : 0);
    C.off_beta = take(8 * Rp);
    C.off_betaf = take(4 * Rp);
    C.off_Eb = take(8 * Rp);
    C.off_Es = take(8 * Rp);
    C.off_red = take(8 * Rp);
    C.off_rep = take(4 * Rp);
    C.off_rung = take(4 * Rp);
    C.off_label = take(4 * Rp);
    C.off_stats = take(5 * 8 * Rp);
    C.off_scr = take(16 * 8 + 16 * 4);
    C.off_sh = take(64);
    return off;
}

// Largest co-resident grid whose slices fit LDS: up to 8 workgroups per CU, never more workgroups
// than one pass over the largest colour class needs.  False when no split fits.
bool coop_plan(const HostInst& H, int block_req, CoopPlan& P) {
    int dev = 0, coop = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop)
        return false;
    const int Rp = H.Rp;
    int block = block_req > 0 ? block_req : 1024;
    block = std::max(block, Rp);
    block = std::min(((block + Rp - 1) / Rp) * Rp, 1024);
    int maxc = 1;
    for (int c = 0; c < H.k; ++c) maxc = std::max(maxc, H.color_ptr[c + 1] - H.color_ptr[c]);
    const int ppb = block / Rp;
    const int need = (maxc + ppb - 1) / ppb;
    const bool uni = Rp % 64 == 0;
    for (int bpc = 8; bpc >= 1; --bpc) {
        const int W = std::max(1, std::min(prop.multiProcessorCount * bpc, need));
        CoopArgs C{};
        const size_t lds = coop_layout(H, W, C);
        if (lds > (size_t)device_lds_limit()) continue;
        int per_cu = 0;
        if (pt_coop_occupancy(H.wm, uni, block, lds, &per_cu) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        if ((int64_t)per_cu * prop.multiProcessorCount < W) continue;
        P.block = block;
        P.grid = W;
        P.lds = lds;
        P.C = C;
        return true;
    }
    return false;
}

int run_coop(const HostInst& H, drift_pt_result* res, CoopPlan P) {
    DevArena A;
    PtInstance I;
    int rc = upload_instance(A, H, I, false, H.flags & DRIFT_PT_RECORD_HISTORY);
    if (rc) return rc;
    I.jbits = nullptr;
    const int Rp = H.Rp;
    GState G;
    PT_HIP(A.alloc(&G.sp, (size_t)H.n * Rp));
    PT_HIP(A.alloc(&G.Eb, 2 * (size_t)Rp));
    PT_HIP(A.alloc(&G.dacc, 3 * (size_t)PT_DACC_SPLIT * Rp));
    PT_HIP(A.alloc(&G.rung, 2 * (size_t)Rp));
    PT_HIP(A.alloc(&G.rep, 2 * (size_t)Rp));
    PT_HIP(A.alloc(&G.label, (size_t)Rp));
    PT_HIP(A.alloc(&G.best, 10));
    PT_HIP(A.alloc(&G.round_base, 1));
    PT_HIP(A.alloc(&P.C.flags, (size_t)P.grid));
    PT_HIP(A.alloc(&P.C.gen, 64));
    PT_HIP(A.alloc(&P.C.abort, 1));
    StreamGuard sg;
    PT_HIP(hipStreamCreate(&sg.s));
    PT_HIP(hipEventCreate(&sg.e0));
    PT_HIP(hipEventCreate(&sg.e1));
    PT_HIP(hipMemsetAsync(G.Eb, 0, 2 * sizeof(int64_t) * Rp, sg.s));
    PT_HIP(hipMemsetAsync(G.dacc, 0, 3 * PT_DACC_SPLIT * sizeof(int64_t) * Rp, sg.s));
    PT_HIP(hipMemsetAsync(G.rung, 0, 2 * sizeof(int32_t) * Rp, sg.s));
    PT_HIP(hipMemsetAsync(G.rep, 0, 2 * sizeof(int32_t) * Rp, sg.s));
    PT_HIP(hipMemsetAsync(G.label, 0, sizeof(int32_t) * Rp, sg.s));
    PT_HIP(hipMemsetAsync(G.best, 0, 10 * sizeof(int64_t), sg.s));
    PT_HIP(hipMemsetAsync(G.round_base, 0, sizeof(int32_t), sg.s));
    PT_HIP(hipMemsetAsync(P.C.flags, 0, sizeof(uint32_t) * P.grid, sg.s));
    PT_HIP(hipMemsetAsync(P.C.gen, 0, sizeof(uint32_t), sg.s));
    PT_HIP(hipMemsetAsync(P.C.abort, 0, sizeof(int32_t), sg.s));
    // initial spins, energies and round-0 state from the global engine's init kernels
    const int iblock = std::min(((std::max(256, Rp) + Rp - 1) / Rp) * Rp, 1024);
    const int reg = std::max(1, std::min(256, (H.n + 1023) / 1024));   // round-end grid
    PT_HIP(pt_launch_g_init(I, G, iblock, reg, sg.s));
    PT_HIP(hipMemsetAsync(G.dacc, 0, 3 * PT_DACC_SPLIT * sizeof(int64_t) * Rp, sg.s));
    P.C.timeout_ticks = barrier_timeout_ticks();
    PT_HIP(hipEventRecord(sg.e0, sg.s));
    if (H.n_rounds > 0)
        PT_HIP(pt_launch_coop(I, G, P.C, P.block, P.grid, P.lds,
                              !(H.flags & DRIFT_PT_PLAIN_LAUNCH), sg.s));
    PT_HIP(hipEventRecord(sg.e1, sg.s));
    PT_HIP(hipStreamSynchronize(sg.s));
    float ms = 0.0f;
    PT_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));
    int32_t aborted = 0;
    PT_HIP(hipMemcpy(&aborted, P.C.abort, sizeof(aborted), hipMemcpyDeviceToHost));
    if (aborted) return fail(DRIFT_PT_EHIP, "coop engine: a grid barrier wait exceeded its bound");
    rc = finish_global(H, I, G, H.n_rounds, res);
    if (rc) return rc;
    res->seconds = ms * 1e-3;
    res->engine_used = DRIFT_PT_ENGINE_COOP;
    res->block_threads_used = P.block;
    res->kernel_seconds = res->seconds;
    res->kernel_attempts = res->attempts;
    res->kernel_launches = 1;
    return 0;
}

// Slabs per replica for the slab engine: about 256 workgroups in all, each slab thick enough
// that its two halo planes on either side stay distinct planes (T + 4 <= L).
int slab_count(const HostInst& H) {
    int NB = std::max(1, std::min(H.lL, (256 + H.R - 1) / H.R));
    while (NB > 1 && (H.lL + NB - 1) / NB + 4 > H.lL) --NB;
    return NB;
}

// power-of-two side (the kernels index by shifts) with at least one quad per plane and colour
bool slab_shape(const HostInst& H) {
    return H.ldim && !(H.lL & (H.lL - 1)) && (H.ldim == 2 ? H.lL >= 8 : H.lL >= 4);
}

bool slab_fits(const HostInst& H) {
    if (!slab_shape(H)) return false;
    const int NB = slab_count(H);
    return pt_slab_lds_bytes(H.n, H.lL, H.ldim, NB) + 32 * 1024 <= (size_t)device_lds_limit();
}

// The persistent slab engine keeps slabs as bits when L is a multiple of 32 (an eighth of the
// byte slab's LDS), else as bytes; whether the grid is co-resident is pslab_block's question.
bool pslab_fits(const HostInst& H) { return slab_shape(H) && (pt_pslab_bits(H.lL) || slab_fits(H)); }

// Results of the slab engines: statistics, best record, final energies and rungs from the
// buffers of parity n_rounds & 1, final spins from spin buffer (n_rounds * S) & 1.
int finish_slab(const HostInst& H, const PtInstance& I, const SState& S, drift_pt_result* res) {
    const int Rp = H.Rp, n = H.n;
    const int fin = H.n_rounds & 1;
    std::vector<int64_t> out((size_t)pt_out_len(H.R));
    PT_HIP(hipMemcpy(out.data(), I.out, out.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t best[5];
    PT_HIP(hipMemcpy(best, S.best + 5 * fin, sizeof(best), hipMemcpyDeviceToHost));
    std::vector<int64_t> E((size_t)Rp);
    std::vector<int32_t> rung((size_t)Rp);
    PT_HIP(hipMemcpy(E.data(), S.Eb + (size_t)fin * Rp, sizeof(int64_t) * Rp, hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(rung.data(), S.rung + (size_t)fin * Rp, sizeof(int32_t) * Rp, hipMemcpyDeviceToHost));
    for (int x = 0; x < 5; ++x) out[x] = best[x];
    out[5] = H.n_rounds;
    for (int r = 0; r < H.R; ++r) {
        pt_out_arr(out.data(), H.R, PT_OUT_E)[r] = E[r];
        pt_out_arr(out.data(), H.R, PT_OUT_RUNG)[r] = rung[r];
    }
    std::vector<int8_t> fs;
    if (H.flags & DRIFT_PT_RECORD_FINAL) {
        const int buf = (int)(((int64_t)H.n_rounds * H.S) & 1);
        std::vector<int8_t> slot_major((size_t)n * Rp);
        PT_HIP(hipMemcpy(slot_major.data(), S.sp + (size_t)buf * Rp * n, slot_major.size(),
                         hipMemcpyDeviceToHost));
        fs.resize((size_t)n * Rp);
        for (int p = 0; p < n; ++p)
            for (int a = 0; a < H.R; ++a) fs[(size_t)p * Rp + a] = slot_major[(size_t)a * n + H.orig[p]];
    }
    return fetch_results(H, I, out, fs, res);
}

// Slab state of one instance: allocations, zeroed bookkeeping, initial spins and energies.
int slab_setup(DevArena& A, const HostInst& H, const PtInstance& I, SState& S, hipStream_t st) {
    const int Rp = H.Rp, n = H.n;
    S.NB = slab_count(H);
    S.upre = 0;
    S.ebar = 0;
    PT_HIP(A.alloc(&S.sp, 2 * (size_t)Rp * n));
    PT_HIP(A.alloc(&S.Eb, 2 * (size_t)Rp));
    PT_HIP(A.alloc(&S.dacc, 3 * (size_t)Rp));
    PT_HIP(A.alloc(&S.einit, (size_t)Rp));
    PT_HIP(A.alloc(&S.rung, 2 * (size_t)Rp));
    PT_HIP(A.alloc(&S.rep, 2 * (size_t)Rp));
    PT_HIP(A.alloc(&S.label, (size_t)Rp));
    PT_HIP(A.alloc(&S.best, 10));
    PT_HIP(A.alloc(&S.round_base, 1));
    PT_HIP(hipMemsetAsync(S.Eb, 0, 2 * sizeof(int64_t) * Rp, st));
    PT_HIP(hipMemsetAsync(S.dacc, 0, 3 * sizeof(int64_t) * Rp, st));
    PT_HIP(hipMemsetAsync(S.einit, 0, sizeof(int64_t) * Rp, st));
    PT_HIP(hipMemsetAsync(S.rung, 0, 2 * sizeof(int32_t) * Rp, st));
    PT_HIP(hipMemsetAsync(S.rep, 0, 2 * sizeof(int32_t) * Rp, st));
    PT_HIP(hipMemsetAsync(S.label, 0, sizeof(int32_t) * Rp, st));
    PT_HIP(hipMemsetAsync(S.best, 0, 10 * sizeof(int64_t), st));
    PT_HIP(hipMemsetAsync(S.round_base, 0, sizeof(int32_t), st));
    PT_HIP(pt_launch_slab_init(I, S, st));
    return 0;
}

int run_slab(const HostInst& H, drift_pt_result* res, int block_req) {
    DevArena A;
    PtInstance I;
    int rc = upload_instance(A, H, I, false, H.flags & DRIFT_PT_RECORD_HISTORY);
    if (rc) return rc;
    const int n = H.n;
    StreamGuard sg;
    PT_HIP(hipStreamCreate(&sg.s));
    PT_HIP(hipEventCreate(&sg.e0));
    PT_HIP(hipEventCreate(&sg.e1));
    SState S;
    rc = slab_setup(A, H, I, S, sg.s);
    if (rc) return rc;
    const size_t lds = pt_slab_lds_bytes(n, H.lL, H.ldim, S.NB);
    PT_HIP(pt_slab_set_lds(H.ldim, S.NB, lds));
    int block = block_req > 0 ? block_req : 512;
    block = std::max(64, std::min(1024, (block / 64) * 64));
    const int per_round = std::max(H.S, 1);
    auto launch_round = [&](int rl) -> int {
        if (H.S == 0) {
            PT_HIP(pt_launch_slab_sweep(I, S, 0, rl, 0, block, lds, sg.s));
            return 0;
        }
        for (int j = 0; j < H.S; ++j) PT_HIP(pt_launch_slab_sweep(I, S, j, rl, 1, block, lds, sg.s));
        return 0;
    };
    GraphGuard gg;
    int chunk = 0;
    if (!(H.flags & DRIFT_PT_NO_GRAPH) && H.n_rounds >= 4) {
        chunk = std::max(1, std::min(H.n_rounds, 1024 / per_round));
        PT_HIP(hipStreamBeginCapture(sg.s, hipStreamCaptureModeThreadLocal));
        for (int rl = 0; rl < chunk; ++rl) {
            rc = launch_round(rl);
            if (rc) {
                hipGraph_t junk = nullptr;
                (void)hipStreamEndCapture(sg.s, &junk);
                if (junk) (void)hipGraphDestroy(junk);
                return rc;
            }
        }
        PT_HIP(pt_launch_slab_tick(S, chunk, sg.s));
        PT_HIP(hipStreamEndCapture(sg.s, &gg.graph));
        PT_HIP(hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0));
    }
    const int replays = chunk ? H.n_rounds / chunk : 0;
    const int rest = H.n_rounds - replays * chunk;
    PT_HIP(hipEventRecord(sg.e0, sg.s));
    for (int g = 0; g < replays; ++g) PT_HIP(hipGraphLaunch(gg.exec, sg.s));
    for (int rl = 0; rl < rest; ++rl) {
        rc = launch_round(rl);
        if (rc) return rc;
    }
    PT_HIP(pt_launch_slab_sweep(I, S, 0, rest, 0, block, lds, sg.s));   // end of the last round
    PT_HIP(hipEventRecord(sg.e1, sg.s));
    PT_HIP(hipStreamSynchronize(sg.s));
    float ms = 0.0f;
    PT_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));

    rc = finish_slab(H, I, S, res);
    if (rc) return rc;
    res->seconds = ms * 1e-3;
    res->engine_used = DRIFT_PT_ENGINE_SLAB;
    res->block_threads_used = block;
    res->kernel_seconds = res->seconds;   // one sweep kernel per sweep carries all the work
    res->kernel_attempts = res->attempts;
    res->kernel_launches = (int64_t)H.n_rounds * per_round + 1;
    return 0;
}

// Block size of the persistent slab engine for this instance, or 0 when its grid cannot be
// co-resident.  upre: the grid also fits with the bit-sliced kernel's buffer of next-sweep
// uniforms (taken when it does; it needs more than one wave per workgroup).
int pslab_block(const HostInst& H, int block_req, size_t& lds, bool* upre = nullptr) {
    int dev = 0, coop = 0, cus = 0;
    if (upre) *upre = false;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const int NB = slab_count(H);
    const size_t lds0 = pt_pslab_lds_bytes(H.lL, H.ldim, NB);
    const size_t lds1 = pt_pslab_bits(H.lL) ? pt_pslab_lds_bytes(H.lL, H.ldim, NB, true) : 0;
    const int64_t W = (int64_t)((H.R + 7) / 8) * 8 * NB;
    auto fits = [&](int block, size_t bytes) -> int {   // 1 fits, 0 does not, -1 query failed
        if (bytes > (size_t)device_lds_limit()) return 0;
        int per_cu = 0;
        if (pt_pslab_occupancy(H.ldim, NB, H.lL, block, bytes, &per_cu) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        return (int64_t)per_cu * cus >= W ? 1 : 0;
    };
    // an explicit block size is taken or refused; otherwise the largest that fits
    const int first = block_req > 0 ? std::max(64, std::min(1024, (block_req / 64) * 64)) : 512;
    for (int block = first; block >= 64; block /= 2) {
        // the uniform path's barrier words are polled by one wave, at most 32 per lane
        if (lds1 && block > 64 && upre && W <= 32 * 64) {
            const int f = fits(block, lds1);
            if (f == 1) {
                lds = lds1;
                *upre = true;
                return block;
            }
        }
        const int f = fits(block, lds0);
        if (f < 0) return 0;
        if (f == 1) {
            lds = lds0;
            return block;
        }
        if (block_req > 0) break;
    }
    return 0;
}

int run_pslab(const HostInst& H, drift_pt_result* res, int block_req);

// The bit-sliced engine's two ways of bringing a round's energy changes to the round end (atomic
// adds before the arrival flags, or the changes inside the arrival and release words) cost
// differently from one device to another: each was measured up to 10% faster than the other on
// different MI355X boards running the same 3-D L=32, R=64 run.  The first run of a grid shape in
// a process therefore times both on a copy of the instance (rounds N and 2N, the difference
// cancelling the fixed costs), twice each in the order 0 1 1 0 so that a clock still ramping up
// favours neither, and keeps the faster minimum.  Both make the same decisions.
int energy_path_choice(const HostInst& H, int block_req, int block) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int, int>, int> cache;
    const auto key = std::make_tuple(H.Rp, H.lL, H.ldim, block, H.S);   // S: barriers per round end
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    HostInst T = H;
    T.flags &= ~(DRIFT_PT_RECORD_HISTORY | DRIFT_PT_RECORD_TRACE | DRIFT_PT_RECORD_FINAL |
                 DRIFT_PT_TIME_KERNELS);
    const double per_round = std::max(1.0, (double)H.R * H.n * std::max(H.S, 1));
    // about 2e9 attempts per timed run, so large instances or many sweeps per round stay short
    const int rounds = (int)std::min(4000.0, std::max(4.0, 2.0e9 / per_round));
    std::vector<int8_t> bs((size_t)std::max(H.n, 1));
    drift_pt_result r{};
    r.best_s = bs.data();
    double secs[2] = {-1.0, -1.0};
    const int order[4] = {0, 1, 1, 0};
    for (int v : order) {
        T.force_ebar = v;
        T.n_rounds = 2;
        clear_result(&r);
        if (run_pslab(T, &r, block_req) != 0) continue;
        T.n_rounds = rounds;
        clear_result(&r);
        if (run_pslab(T, &r, block_req) != 0) continue;
        const double t1 = r.seconds;
        T.n_rounds = 2 * rounds;
        clear_result(&r);
        if (run_pslab(T, &r, block_req) != 0) continue;
        const double d = r.seconds - t1 > 0.0 ? r.seconds - t1 : 0.5 * r.seconds;
        if (secs[v] < 0.0 || d < secs[v]) secs[v] = d;
    }
    (void)hipGetLastError();
    g_err.clear();
    const int choice = secs[1] > 0.0 && (secs[0] <= 0.0 || secs[1] < secs[0]) ? 1 : 0;
    if (secs[0] <= 0.0 && secs[1] <= 0.0) return choice;   // neither trial ran: nothing learned
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = choice;
    return choice;
}

// Bit slabs run on the neighbour-synchronised kernel unless the run records the best energy of
// every round or the rung trace (both need the whole ladder each round) or
// DRIFT_PT_PSLAB_SYNC=grid asks for the grid-barrier kernel.
bool pslab_use_nbr(const HostInst& H) {
    if (!pt_pslab_bits(H.lL)) return false;
    if (H.flags & (DRIFT_PT_RECORD_HISTORY | DRIFT_PT_RECORD_TRACE)) return false;
    const char* e = std::getenv("DRIFT_PT_PSLAB_SYNC");
    return !(e && std::strcmp(e, "grid") == 0);
}

// Colour-list position of lattice site `site` (the parity colouring in index order).
int slab_pos_of(int site, const HostInst& H) {
    const int L = H.lL;
    const int x = site % L, y = (site / L) % L, z = H.ldim == 3 ? site / (L * L) : 0;
    const int par = (H.ldim == 3 ? x + y + z : x + (site / L)) & 1;
    return par * (H.n / 2) + (site >> 1);
}

// The neighbour-synchronised kernel leaves one best record per slot: the run's best is their
// lexicographic minimum over (energy, round, rung), its state that slot's bits.
int nbr_best(const HostInst& H, const PtInstance& I, const SState& S, const NbrArgs& N) {
    std::vector<int64_t> sb(4 * (size_t)H.R);
    PT_HIP(hipMemcpy(sb.data(), N.sbest, sizeof(int64_t) * sb.size(), hipMemcpyDeviceToHost));
    int w = 0;
    for (int x = 1; x < H.R; ++x) {
        const int64_t* c = &sb[4 * (size_t)x];
        const int64_t* m = &sb[4 * (size_t)w];
        if (c[0] < m[0] || (c[0] == m[0] && (c[1] < m[1] || (c[1] == m[1] && c[2] < m[2])))) w = x;
    }
    const int64_t rec[5] = {sb[4 * (size_t)w], sb[4 * (size_t)w + 1], sb[4 * (size_t)w + 2], w, 1};
    PT_HIP(hipMemcpy(S.best + 5 * (H.n_rounds & 1), rec, sizeof(rec), hipMemcpyHostToDevice));
    const size_t words = (size_t)H.n / 32;
    std::vector<uint32_t> bits(words);
    PT_HIP(hipMemcpy(bits.data(), N.bestbits + (size_t)w * words, sizeof(uint32_t) * words,
                     hipMemcpyDeviceToHost));
    std::vector<int8_t> st((size_t)H.n);
    for (int site = 0; site < H.n; ++site)
        st[(size_t)slab_pos_of(site, H)] = ((bits[(size_t)site >> 5] >> (site & 31)) & 1u) ? -1 : 1;
    PT_HIP(hipMemcpy(I.best_s, st.data(), st.size(), hipMemcpyHostToDevice));
    return 0;
}

int run_pslab(const HostInst& H, drift_pt_result* res, int block_req) {
    size_t lds = 0;
    bool upre = false;
    const int block = pslab_block(H, block_req, lds, &upre);
    if (!block) return fail(DRIFT_PT_EINVAL, "pslab engine: the slab grid cannot be co-resident");
    const bool nbr = pslab_use_nbr(H);
    // chosen before this run allocates anything; DRIFT_PT_PSLAB_ENERGY_PATH=atomic|words fixes it
    int ebar = 0;
    if (upre && !nbr) {
        const char* ep = std::getenv("DRIFT_PT_PSLAB_ENERGY_PATH");
        if (H.force_ebar >= 0) ebar = H.force_ebar;
        else if (ep && std::strcmp(ep, "atomic") == 0) ebar = 0;
        else if (ep && std::strcmp(ep, "words") == 0) ebar = 1;
        else ebar = energy_path_choice(H, block_req, block);
    }
    DevArena A;
    PtInstance I;
    int rc = upload_instance(A, H, I, false, H.flags & DRIFT_PT_RECORD_HISTORY);
    if (rc) return rc;
    StreamGuard sg;
    PT_HIP(hipStreamCreate(&sg.s));
    PT_HIP(hipEventCreate(&sg.e0));
    PT_HIP(hipEventCreate(&sg.e1));
    SState S;
    rc = slab_setup(A, H, I, S, sg.s);
    if (rc) return rc;
    S.hbits = nullptr;
    S.upre = upre ? 1 : 0;
    S.ebar = ebar;
    if (pt_pslab_bits(H.lL)) {
        const size_t words = 2 * (size_t)H.Rp * (size_t)H.n / 32;
        PT_HIP(A.alloc(&S.hbits, words));
        PT_HIP(hipMemsetAsync(S.hbits, 0, words * sizeof(uint32_t), sg.s));
    }
    CoopArgs C{};
    const int W = pt_pslab_grid(I, S);
    PT_HIP(A.alloc(&C.flags, (size_t)W));
    PT_HIP(A.alloc(&C.gen, 64));
    PT_HIP(A.alloc(&C.abort, 1));
    PT_HIP(hipMemsetAsync(C.flags, 0, sizeof(uint32_t) * W, sg.s));
    PT_HIP(hipMemsetAsync(C.gen, 0, sizeof(uint32_t), sg.s));
    PT_HIP(hipMemsetAsync(C.abort, 0, sizeof(int32_t), sg.s));
    if (S.ebar) {   // arrival and release words that carry the energy changes (tags start at 0)
        PT_HIP(A.alloc(&C.eflags, (size_t)W));
        PT_HIP(A.alloc(&C.erel, 64));
        PT_HIP(hipMemsetAsync(C.eflags, 0, sizeof(uint64_t) * W, sg.s));
        PT_HIP(hipMemsetAsync(C.erel, 0, sizeof(uint64_t) * 64, sg.s));
    }
    C.timeout_ticks = barrier_timeout_ticks();
    NbrArgs N{};
    if (nbr) {
        const size_t nbox = (size_t)NBR_ROUNDS * H.Rp * S.NB;
        PT_HIP(A.alloc(&N.box, nbox));
        PT_HIP(A.alloc(&N.halo, 2 * (size_t)H.Rp * H.n / 32));
        PT_HIP(A.alloc(&N.sbest, 4 * (size_t)H.Rp));
        PT_HIP(A.alloc(&N.bestbits, (size_t)H.Rp * H.n / 32));
        PT_HIP(hipMemsetAsync(N.box, 0xFF, sizeof(uint64_t) * nbox, sg.s));
        PT_HIP(hipMemsetAsync(N.halo, 0xFF, sizeof(uint64_t) * 2 * H.Rp * H.n / 32, sg.s));
        // rung and pair counters are added into
        PT_HIP(hipMemsetAsync(I.out, 0, sizeof(int64_t) * pt_out_len(H.R), sg.s));
    }
    PT_HIP(hipEventRecord(sg.e0, sg.s));
    if (nbr)
        PT_HIP(pt_launch_pslab_nbr(I, S, C, N, block, lds, !(H.flags & DRIFT_PT_PLAIN_LAUNCH), sg.s));
    else
        PT_HIP(pt_launch_pslab(I, S, C, block, lds, !(H.flags & DRIFT_PT_PLAIN_LAUNCH), sg.s));
    PT_HIP(hipEventRecord(sg.e1, sg.s));
    PT_HIP(hipStreamSynchronize(sg.s));
    float ms = 0.0f;
    PT_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));
    int32_t aborted = 0;
    PT_HIP(hipMemcpy(&aborted, C.abort, sizeof(aborted), hipMemcpyDeviceToHost));
    if (aborted)
        return fail(DRIFT_PT_EHIP, nbr ? "pslab engine: a neighbour wait exceeded its bound"
                                       : "pslab engine: a grid barrier wait exceeded its bound");
    if (nbr && (rc = nbr_best(H, I, S, N)) != 0) return rc;
    rc = finish_slab(H, I, S, res);
    if (rc) return rc;
    res->seconds = ms * 1e-3;
    res->engine_used = DRIFT_PT_ENGINE_PSLAB;
    res->block_threads_used = block;
    res->kernel_seconds = res->seconds;
    res->kernel_attempts = res->attempts;
    res->kernel_launches = 1;
    return 0;
}

// LDS engine fit: the dynamic area plus ~1 KB of the kernel's static shared variables.
bool fits_lds(const HostInst& H) {
    return pt_lds_bytes(H.n, H.Rp, H.k) + 1024 <= (size_t)device_lds_limit();
}

// Run one prepared instance on a resolved engine (not AUTO).
int run_engine(const HostInst& H, int engine, int block, drift_pt_result* res) {
    if (H.has_target && engine != DRIFT_PT_ENGINE_LDS && engine != DRIFT_PT_ENGINE_GLOBAL &&
        engine != DRIFT_PT_ENGINE_LATTICE && engine != DRIFT_PT_ENGINE_SPLIT)
        return fail(DRIFT_PT_EINVAL, "DRIFT_PT_STOP_AT_TARGET needs the lds, split, global or "
                                     "lattice engine (the other persistent engines run every round "
                                     "in one launch)");
    switch (engine) {
    case DRIFT_PT_ENGINE_SLAB:
        if (!slab_fits(H))
            return fail(DRIFT_PT_EINVAL, "slab engine needs a periodic L^2/L^3 lattice (L even >= 4) "
                                         "with +/-1 couplings, no field, the parity colouring, and "
                                         "slabs that fit LDS");
        return run_slab(H, res, block);
    case DRIFT_PT_ENGINE_LATTICE:
        if (!H.ldim)
            return fail(DRIFT_PT_EINVAL, "lattice engine needs a periodic L^2/L^3 lattice (L even >= 4) "
                                         "with +/-1 couplings, no field and the parity colouring");
        return run_global(H, res, block, true);
    case DRIFT_PT_ENGINE_LDS: {
        if (!fits_lds(H)) return fail(DRIFT_PT_EINVAL, "instance does not fit one workgroup's LDS");
        std::vector<HostInst> hs{H};
        return run_lds(hs, {res}, block);
    }
    case DRIFT_PT_ENGINE_PSLAB:
        if (!pslab_fits(H))
            return fail(DRIFT_PT_EINVAL, "pslab engine needs a periodic L^2/L^3 lattice (L a power of "
                                         "two) with +/-1 couplings, no field, the parity colouring, "
                                         "and slabs that fit LDS");
        return run_pslab(H, res, block);
    case DRIFT_PT_ENGINE_COOP: {
        CoopPlan P;
        if (!coop_plan(H, block, P))
            return fail(DRIFT_PT_EINVAL, "coop engine: no co-resident grid whose row slices fit LDS");
        return run_coop(H, res, P);
    }
    case DRIFT_PT_ENGINE_GLOBAL:
        return run_global(H, res, block, false);
    case DRIFT_PT_ENGINE_SPLIT: {
        if (!split_fits(H))
            return fail(DRIFT_PT_EINVAL, "split engine: one replica's spins do not fit one "
                                         "workgroup's LDS");
        std::vector<HostInst> hs{H};
        return run_split(hs, {res}, block);
    }
    default:
        return fail(DRIFT_PT_EINVAL, "unknown engine");
    }
}

// The engine AUTO takes without tuning.  A single instance on the LDS engine runs on one CU
// (measured 0.25-0.4 attempts/ns), while the multi-CU engines pay a fixed 5-30 us per round, so the
// LDS engine only takes rounds of at most 8192 attempts.  Lattices go to the persistent slab
// engine; other instances to the coop engine (one launch, a grid barrier per colour) up to 2^17
// attempts per sweep and to the per-launch global engine above (ER N=10^4, R=32: 6.7 s vs 8.1 s
// for 2e5 rounds; C3: 56 vs 49 attempts/ns).
// Parts per one-slot group for one instance on its own, by default: the largest power of two p
// (at most 64 and the largest colour class's word count) whose R p workgroups of 1024 threads are
// all resident, when the instance has n >= 8000 positions, at most four colour classes and fewer
// replicas than compute units; else 0.  Each traded class costs a hop per sweep: on 3-regular
// graphs (four classes) and 3-D glasses (two) of 8000 to 131 072 spins at R = 16 and 32 the
// groups in parts ran 6-57% faster than the engine otherwise chosen, on Erdos-Renyi graphs (five
// to seven classes) 9-22% slower.
int parts_default(const HostInst& H) {
    if (H.n < 8000 || H.k > 4 || H.R >= device_cus() || split_rg_max(H) < 1) return 0;
    const size_t lds = split_lds(H, 1, false, true);
    int per_cu = 0;
    if (lds > (size_t)device_lds_limit() ||
        pt_split_occupancy(H.wm, 1024, lds, &per_cu, false, true) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    const int64_t cap = std::min<int64_t>((int64_t)per_cu * device_cus(), 1024);
    int np = 1;
    while (np < 64 && (int64_t)H.R * np * 2 <= cap && np * 2 <= class_words_max(H)) np *= 2;
    return np >= 2 ? np : 0;
}

int default_engine(const HostInst& H) {
    const int64_t per_sweep = (int64_t)H.R * H.n;
    if (fits_lds(H) && per_sweep <= 8192) return DRIFT_PT_ENGINE_LDS;
    // (not when the caller fixed the split engine's launch shape: then AUTO keeps its other rules)
    const bool parts = H.rg_req == 0 && H.np_req == 0 && parts_default(H) >= 2;
    if (H.has_target)
        return parts ? DRIFT_PT_ENGINE_SPLIT
                     : (H.ldim ? DRIFT_PT_ENGINE_LATTICE : DRIFT_PT_ENGINE_GLOBAL);
    if (H.ldim && pslab_fits(H)) {
        size_t lds = 0;
        if (pslab_block(H, 0, lds)) return DRIFT_PT_ENGINE_PSLAB;
    }
    if (H.ldim && slab_fits(H)) return DRIFT_PT_ENGINE_SLAB;
    if (parts) return DRIFT_PT_ENGINE_SPLIT;
    CoopPlan P;
    if (per_sweep <= (1 << 17) && coop_plan(H, 0, P)) return DRIFT_PT_ENGINE_COOP;
    return H.ldim ? DRIFT_PT_ENGINE_LATTICE : DRIFT_PT_ENGINE_GLOBAL;
}

// The launch shape default_engine means when it picks the split engine for one instance: one
// slot per group in parts_default parts (unless the caller fixed either).
void default_split_shape(HostInst& H) {
    if (H.rg_req > 0 || H.np_req > 0) return;
    const int np = parts_default(H);
    if (np >= 2) {
        H.rg_req = 1;
        H.np_req = np;
    }
}

// ------------------------------------------------------------------------------------------
// autotuner: timed trials per shape class, cached for the process
// ------------------------------------------------------------------------------------------
using TuneKey = std::tuple<int, int, int, int, int, int, int, int, int, int>;

struct TuneEntry {
    int engine = 0, block = 0, rg = 0, cl = 0, np = 0;
};

std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;

// launch shapes chosen by the batch autotuner (tune_batch), by batch shape class
using BatchKey = std::tuple<int, int, int, int, int, int, int, int, int>;

struct BatchShape {
    int block = 0, rg_cap = 0;   // rg_cap: split engine, slots per group min(rg_cap, largest fit)
    int cl = 0;                  // split engine: couplings in LDS
};

std::mutex g_btune_mu;
std::map<BatchKey, BatchShape> g_btune;

TuneKey tune_key(const HostInst& H) {
    int nb = 0;
    while ((1 << nb) < H.n) ++nb;                               // ceil(log2 n)
    const int half_degree = H.n ? (int)std::lround(2.0 * H.nnz / H.n) : 0;
    // whether the couplings fit 16-bit LDS records decides which candidates exist, and so does
    // a target (only the engines that can stop at one are candidates)
    return TuneKey(H.wm, H.Rp, H.ldim, H.lL, nb, std::min(half_degree, 256), H.S, H.has_h ? 1 : 0,
                   H.cl_ok ? 1 : 0, H.has_target ? 1 : 0);
}

// Timed trials of every engine and launch shape that can run H.  The first candidate is what
// default_engine takes with its own launch shape (block 0); the choice stays there unless another
// candidate is more than 3% faster.  Every candidate is timed in two passes, the second in the
// opposite order, and keeps the faster of its two times, so drift of the device's clock over the
// trials favours no candidate.
int tune(const HostInst& H, int trial_rounds, drift_pt_tuning* table, TuneEntry& choice) {
    struct Cand {
        int engine, block, rg, cl, np;
        double secs;
    };
    std::vector<Cand> cands;
    {
        // the default, with the launch shape it means (a split default: one-slot groups in parts)
        HostInst D = H;
        const int e = default_engine(D);
        if (e == DRIFT_PT_ENGINE_SPLIT) default_split_shape(D);
        cands.push_back({e, 0, e == DRIFT_PT_ENGINE_SPLIT ? D.rg_req : 0, 0,
                         e == DRIFT_PT_ENGINE_SPLIT ? D.np_req : 0, -1.0});
    }
    const int blocks[3] = {256, 512, 1024};
    for (int b : blocks) {
        if (fits_lds(H)) cands.push_back({DRIFT_PT_ENGINE_LDS, b, 0, 0, 0, -1.0});
        if (H.ldim && slab_fits(H)) cands.push_back({DRIFT_PT_ENGINE_SLAB, b, 0, 0, 0, -1.0});
        size_t plds = 0;
        if (H.ldim && pslab_fits(H) && pslab_block(H, b, plds) == b)
            cands.push_back({DRIFT_PT_ENGINE_PSLAB, b, 0, 0, 0, -1.0});
        if (H.ldim) cands.push_back({DRIFT_PT_ENGINE_LATTICE, b, 0, 0, 0, -1.0});
        cands.push_back({DRIFT_PT_ENGINE_GLOBAL, b, 0, 0, 0, -1.0});
        CoopPlan P;
        if (coop_plan(H, b, P)) cands.push_back({DRIFT_PT_ENGINE_COOP, b, 0, 0, 0, -1.0});
        // the split engine, every group size whose groups can all be resident, with the
        // couplings in HBM and (where they qualify) in LDS
        for (int cl = 0; cl <= (H.cl_ok ? 1 : 0); ++cl)
            for (int rg = split_rg_max(H, cl); rg >= 1; rg >>= 1) {
                const size_t lds = split_lds(H, rg, cl);
                int per_cu = 0;
                if (pt_split_occupancy(H.wm, b, lds, &per_cu, cl) != hipSuccess) {
                    (void)hipGetLastError();
                    continue;
                }
                if ((int64_t)per_cu * device_cus() >= (H.R + rg - 1) / rg)
                    cands.push_back({DRIFT_PT_ENGINE_SPLIT, b, rg, cl, 1, -1.0});
            }
        // one-slot groups split into parts (the packed spin-word trade), when the groups alone
        // leave compute units idle: every power of two of parts whose workgroups can all be
        // resident, at 1024 threads (fewer threads ran slower at every count measured)
        if (b == 1024 && split_rg_max(H) >= 1 && H.R < device_cus()) {
            const size_t lds = split_lds(H, 1, false, true);
            int per_cu = 0;
            if (lds <= (size_t)device_lds_limit() &&
                pt_split_occupancy(H.wm, b, lds, &per_cu, false, true) == hipSuccess) {
                const int64_t cap = (int64_t)per_cu * device_cus();
                for (int np = 2; np <= 64 && np <= class_words_max(H) && (int64_t)H.R * np <= cap &&
                                 (int64_t)H.R * np <= 1024;
                     np *= 2)
                    cands.push_back({DRIFT_PT_ENGINE_SPLIT, b, 1, 0, np, -1.0});
            } else {
                (void)hipGetLastError();
            }
        }
    }
    if (H.has_target) {
        // a run that stops at a target: only the engines that can
        std::vector<Cand> keep;
        for (const Cand& c : cands)
            if (c.engine == DRIFT_PT_ENGINE_LDS || c.engine == DRIFT_PT_ENGINE_GLOBAL ||
                c.engine == DRIFT_PT_ENGINE_LATTICE || c.engine == DRIFT_PT_ENGINE_SPLIT)
                keep.push_back(c);
        cands.swap(keep);
    }
    const double per_round = std::max(1.0, (double)H.R * H.n * std::max(H.S, 1));
    const int rounds = trial_rounds > 0 ? trial_rounds
                                        : (int)std::min(2000.0, std::max(4.0, 2.5e8 / per_round));
    HostInst T = H;
    T.flags &= ~(DRIFT_PT_RECORD_HISTORY | DRIFT_PT_RECORD_TRACE | DRIFT_PT_RECORD_FINAL |
                 DRIFT_PT_TIME_KERNELS | DRIFT_PT_STOP_AT_TARGET);
    T.has_target = 0;   // trials run every round
    std::vector<int8_t> bs((size_t)std::max(H.n, 1));
    drift_pt_result r{};
    r.best_s = bs.data();
    // Each trial runs `rounds` and then 2 * `rounds` rounds; the difference is the time of
    // `rounds` rounds with the fixed costs (staging, ramp, launch) cancelled, which is what a long
    // run pays.
    auto trial = [&](Cand& c) {
        double secs = -1.0;
        T.rg_req = c.rg;
        T.np_req = c.np;
        T.flags = c.cl ? (T.flags | DRIFT_PT_COUPLINGS_LDS) : (T.flags & ~DRIFT_PT_COUPLINGS_LDS);
        T.n_rounds = 2;                                      // warm-up: code objects, first touch
        clear_result(&r);
        if (run_engine(T, c.engine, c.block, &r) == 0) {
            T.n_rounds = rounds;
            clear_result(&r);
            if (run_engine(T, c.engine, c.block, &r) == 0) {
                const double t1 = r.seconds;
                T.n_rounds = 2 * rounds;
                clear_result(&r);
                if (run_engine(T, c.engine, c.block, &r) == 0)
                    secs = r.seconds - t1 > 0.0 ? r.seconds - t1 : 0.5 * r.seconds;
            }
        }
        (void)hipGetLastError();
        if (secs > 0.0 && (c.secs < 0.0 || secs < c.secs)) c.secs = secs;
    };
    for (Cand& c : cands) trial(c);
    for (size_t i = cands.size(); i-- > 0;) trial(cands[i]);
    g_err.clear();
    int best = -1;
    for (size_t i = 0; i < cands.size(); ++i)
        if (cands[i].secs > 0.0 && (best < 0 || cands[i].secs < cands[best].secs)) best = (int)i;
    if (best < 0) return fail(DRIFT_PT_EHIP, "autotune: no candidate engine ran");
    if (cands[0].secs > 0.0 && cands[best].secs > cands[0].secs / 1.03) best = 0;
    choice.engine = cands[best].engine;
    choice.block = cands[best].block;
    choice.rg = cands[best].rg;
    choice.cl = cands[best].cl;
    choice.np = cands[best].np;
    if (table) {
        std::memset(table, 0, sizeof(*table));
        table->trial_rounds = rounds;
        int slot = 0;
        for (const Cand& c : cands) {
            if (slot >= DRIFT_PT_TUNE_MAX) break;
            table->cand_engine[slot] = c.engine;
            table->cand_block[slot] = c.block;
            table->cand_group[slot] = c.rg;
            table->cand_lds_couplings[slot] = c.cl;
            table->cand_parts[slot] = c.np;
            table->cand_seconds[slot] = c.secs;
            ++slot;
        }
        table->n_candidates = slot;
        table->engine = choice.engine;
        table->block_threads = choice.block;
        table->slots_per_group = choice.rg;
        table->couplings_in_lds = choice.cl;
        table->parts_per_group = choice.np;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[tune_key(H)] = choice;
    return 0;
}

}  // namespace

extern "C" {

int drift_pt_abi_version(void) { return DRIFT_PT_ABI_VERSION; }

const char* drift_pt_last_error(void) { return g_err.c_str(); }

int drift_pt_device_available(void) {
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) {
        (void)hipGetLastError();
        return 0;
    }
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int drift_pt_greedy_coloring(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                             int32_t* color_ptr, int32_t* color_spins) {
    if (n < 0 || !row_ptr || !color_ptr || (n > 0 && (!color_spins || !col_idx)))
        return fail(DRIFT_PT_EINVAL, "bad arguments to drift_pt_greedy_coloring");
    if (n == 0) {
        color_ptr[0] = 0;
        color_ptr[1] = 0;
        return 1;
    }
    for (int i = 0; i < n; ++i)
        for (int t = row_ptr[i]; t < row_ptr[i + 1]; ++t)
            if (col_idx[t] < 0 || col_idx[t] >= n) return fail(DRIFT_PT_EINVAL, "col_idx out of range");
    return greedy_color(n, row_ptr, col_idx, color_ptr, color_spins);
}

int drift_pt_run(const drift_pt_problem* problem, const drift_pt_params* params,
                 drift_pt_result* result) {
    g_err.clear();
    if (!result || !result->best_s) return fail(DRIFT_PT_EINVAL, "result.best_s is required");
    clear_result(result);
    HostInst H;
    int rc = prepare(problem, params, H);
    if (rc) return rc;
    rc = check_device();
    if (rc) return rc;
    detect_lattice(problem, H);
    int engine = params->engine, block = params->block_threads;
    if (engine == DRIFT_PT_ENGINE_AUTO) {
        engine = default_engine(H);
        if (!(params->flags & DRIFT_PT_AUTOTUNE) && engine == DRIFT_PT_ENGINE_SPLIT)
            default_split_shape(H);
        if (params->flags & DRIFT_PT_AUTOTUNE) {
            TuneEntry choice;
            bool hit = false;
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tune.find(tune_key(H));
                if (it != g_tune.end()) {
                    choice = it->second;
                    hit = true;
                }
            }
            if (!hit) {
                rc = tune(H, 0, nullptr, choice);
                if (rc) return rc;
            }
            engine = choice.engine;
            block = choice.block;
            H.rg_req = choice.rg;
            H.np_req = choice.np;
            if (choice.cl) H.flags |= DRIFT_PT_COUPLINGS_LDS;
        }
    }
    return run_engine(H, engine, block, result);
}

int drift_pt_autotune(const drift_pt_problem* problem, const drift_pt_params* params,
                      int32_t trial_rounds, drift_pt_tuning* out) {
    g_err.clear();
    if (!out || !params) return fail(DRIFT_PT_EINVAL, "autotune needs params and an output table");
    HostInst H;
    int rc = prepare(problem, params, H);
    if (rc) return rc;
    rc = check_device();
    if (rc) return rc;
    detect_lattice(problem, H);
    TuneEntry choice;
    return tune(H, trial_rounds, out, choice);
}

void drift_pt_autotune_clear(void) {
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune.clear();
    }
    std::lock_guard<std::mutex> lk(g_btune_mu);
    g_btune.clear();
}

}  // extern "C"

namespace {

// Preparation shared by the two batch entry points: every instance validated, coloured,
// relabelled and given its thresholds; all must fit one workgroup's LDS.
int prepare_batch(const drift_pt_problem* problems, const drift_pt_params* params,
                  int32_t n_instances, drift_pt_result* results, std::vector<HostInst>& hs,
                  std::vector<drift_pt_result*>& rs, bool require_lds) {
    if (n_instances < 1 || !problems || !params || !results)
        return fail(DRIFT_PT_EINVAL, "bad batch arguments");
    hs.assign((size_t)n_instances, HostInst());
    rs.assign((size_t)n_instances, nullptr);
    for (int i = 0; i < n_instances; ++i)
        if (!results[i].best_s) return fail(DRIFT_PT_EINVAL, "result.best_s is required");
    // instances are prepared independently (colouring, relabelling, thresholds), on several host
    // threads for large batches; errors are reported for the first failing instance
    std::vector<int> rcs((size_t)n_instances, 0);
    std::vector<std::string> errs((size_t)n_instances);
    auto work = [&](int first, int stride) {
        for (int i = first; i < n_instances; i += stride) {
            clear_result(&results[i]);
            rcs[i] = prepare(&problems[i], &params[i], hs[i]);
            if (rcs[i]) errs[i] = g_err;
        }
    };
    const int nthr = n_instances >= 64
                         ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()))
                         : 1;
    if (nthr > 1) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthr; ++t) pool.emplace_back(work, t, nthr);
        for (std::thread& th : pool) th.join();
    } else {
        work(0, 1);
    }
    for (int i = 0; i < n_instances; ++i) {
        if (rcs[i]) return fail(rcs[i], "instance " + std::to_string(i) + ": " + errs[i]);
        rs[i] = &results[i];
    }
    int rc = check_device();
    if (rc) return rc;
    for (int i = 0; i < n_instances; ++i) {
        HostInst& H = hs[i];
        if (require_lds && !fits_lds(H))
            return fail(DRIFT_PT_EINVAL, "batched instances must fit one workgroup's LDS");
        if (!fits_lds(H)) detect_lattice(&problems[i], H);
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// batch autotuner: the launch shape of each persistent launch of a batch, timed on the batch
// ------------------------------------------------------------------------------------------

// Shape class of a batch: engine (1 LDS, 7 split), weight type, largest padded slot count,
// ceil(log2) of the largest n and of the instance count, mean degree, sweeps per round, field,
// and whether every instance's couplings fit 16-bit LDS records.
BatchKey batch_key(int engine, const std::vector<HostInst>& hs) {
    int maxRp = 0, maxn = 1, S = 0, hh = 0, cl = 1;
    double deg = 0.0, n_all = 0.0;
    for (const HostInst& H : hs) {
        maxRp = std::max(maxRp, H.Rp);
        maxn = std::max(maxn, H.n);
        S = std::max(S, H.S);
        hh |= H.has_h ? 1 : 0;
        cl &= H.cl_ok ? 1 : 0;
        deg += H.nnz;
        n_all += H.n;
    }
    auto clog2 = [](int64_t v) {
        int b = 0;
        while (((int64_t)1 << b) < v) ++b;
        return b;
    };
    return BatchKey(engine, hs.empty() ? 0 : hs[0].wm, maxRp, clog2(maxn), clog2((int64_t)hs.size()),
                    (int)std::lround(n_all > 0 ? deg / n_all : 0.0), S, hh, cl);
}

int run_split(const std::vector<HostInst>& hs, const std::vector<drift_pt_result*>& results,
              int block_req);

// Timed trials of launch shapes for one persistent launch of a batch (the LDS engine: block
// sizes; the split engine: group-size caps x block sizes).  Every trial runs the whole batch for
// T and then 2T rounds (no records, no target) and takes the difference, so fixed costs cancel;
// each candidate is timed in two passes of opposite order and keeps its faster time.  The first
// candidate is the library's default shape, which stays unless another is more than 3% faster.
int tune_batch(int engine, const std::vector<HostInst>& hs, BatchShape& out) {
    const BatchKey key = batch_key(engine, hs);
    {
        std::lock_guard<std::mutex> lk(g_btune_mu);
        auto it = g_btune.find(key);
        if (it != g_btune.end()) {
            out = it->second;
            return 0;
        }
    }
    struct Cand {
        BatchShape shape;
        double secs;
    };
    std::vector<Cand> cands;
    cands.push_back({BatchShape{}, -1.0});
    int maxRp = 1;
    for (const HostInst& H : hs) maxRp = std::max(maxRp, H.Rp);
    const int blocks[3] = {256, 512, 1024};
    if (engine == DRIFT_PT_ENGINE_LDS) {
        for (int b : blocks) {
            bool ok = b >= maxRp;
            for (const HostInst& H : hs) ok = ok && b % H.Rp == 0;
            if (ok) cands.push_back({BatchShape{b, 0}, -1.0});
        }
    } else {
        bool cl_ok = true;
        for (const HostInst& H : hs) cl_ok = cl_ok && H.cl_ok;
        for (int cl = 0; cl <= (cl_ok ? 1 : 0); ++cl) {
            int top = 1;
            for (const HostInst& H : hs) top = std::max(top, split_rg_max(H, cl));
            for (int c = top; c >= 1; c >>= 1)
                for (int b : blocks) cands.push_back({BatchShape{b, c, cl}, -1.0});
        }
    }
    int64_t per_round = 0;
    int max_rounds = 0;
    for (const HostInst& H : hs) {
        per_round += (int64_t)H.R * H.n * std::max(H.S, 1);
        max_rounds = std::max(max_rounds, H.n_rounds);
    }
    // about 5e8 attempts per trial run, at least two rounds
    const int T = (int)std::max<int64_t>(2, std::min<int64_t>(64, (int64_t)5e8 / std::max<int64_t>(per_round, 1)));
    std::vector<HostInst> trial = hs;
    std::vector<std::vector<int8_t>> bs(hs.size());
    std::vector<drift_pt_result> res(hs.size());
    std::vector<drift_pt_result*> rp(hs.size());
    for (size_t i = 0; i < hs.size(); ++i) {
        trial[i].flags &= ~(DRIFT_PT_RECORD_HISTORY | DRIFT_PT_RECORD_TRACE | DRIFT_PT_RECORD_FINAL |
                            DRIFT_PT_TIME_KERNELS);
        trial[i].has_target = 0;
        bs[i].resize((size_t)std::max(hs[i].n, 1));
        std::memset(&res[i], 0, sizeof(res[i]));
        res[i].best_s = bs[i].data();
        rp[i] = &res[i];
    }
    auto timed = [&](const BatchShape& sh, int rounds, double& secs) {
        for (HostInst& H : trial) {
            H.n_rounds = rounds;
            if (engine == DRIFT_PT_ENGINE_SPLIT) {
                H.rg_req = sh.rg_cap > 0 ? std::max(1, std::min(sh.rg_cap, split_rg_max(H, sh.cl))) : 0;
                H.flags = sh.cl ? (H.flags | DRIFT_PT_COUPLINGS_LDS) : (H.flags & ~DRIFT_PT_COUPLINGS_LDS);
            }
        }
        const int rc = engine == DRIFT_PT_ENGINE_LDS ? run_lds(trial, rp, sh.block)
                                                     : run_split(trial, rp, sh.block);
        (void)hipGetLastError();
        secs = rc == 0 ? res[0].seconds : -1.0;
        return rc == 0;
    };
    auto one = [&](Cand& c) {
        double t1 = 0.0, t2 = 0.0;
        if (timed(c.shape, T, t1) && timed(c.shape, 2 * T, t2)) {
            const double d = t2 - t1 > 0.0 ? t2 - t1 : 0.5 * t2;
            if (c.secs < 0.0 || d < c.secs) c.secs = d;
        }
    };
    for (Cand& c : cands) one(c);
    for (size_t i = cands.size(); i-- > 0;) one(cands[i]);
    g_err.clear();
    int best = -1;
    for (size_t i = 0; i < cands.size(); ++i)
        if (cands[i].secs > 0.0 && (best < 0 || cands[i].secs < cands[best].secs)) best = (int)i;
    if (best < 0) return fail(DRIFT_PT_EHIP, "batch autotune: no launch shape ran");
    if (cands[0].secs > 0.0 && cands[best].secs > cands[0].secs / 1.03) best = 0;
    out = cands[best].shape;
    std::lock_guard<std::mutex> lk(g_btune_mu);
    g_btune[key] = out;
    return 0;
}

// The batch: instances that fit one workgroup's LDS in one persistent launch per weight type;
// larger ones whose single replica fits on the split engine, in one persistent launch per weight
// type, when their groups number at least one per compute unit; the rest on the per-launch
// engines, four at a time on streams of their own.  A batch shares one workgroup size per launch:
// the largest requested (0: the library's choice).
int run_batch(std::vector<HostInst>& hs, const std::vector<drift_pt_result*>& rs) {
    std::map<int, std::vector<size_t>> by_wm, split_wm;
    std::vector<size_t> big;
    for (size_t i = 0; i < hs.size(); ++i) {
        if (fits_lds(hs[i])) by_wm[hs[i].wm].push_back(i);
        else big.push_back(i);
    }
    auto part_of = [&](const std::vector<size_t>& idx, std::vector<HostInst>& part,
                       std::vector<drift_pt_result*>& prs, int& block) {
        block = 0;
        for (size_t i : idx) {
            part.push_back(hs[i]);
            prs.push_back(rs[i]);
            block = std::max(block, hs[i].block_req);
        }
    };
    bool tune = false;
    for (const HostInst& H : hs) tune = tune || (H.flags & DRIFT_PT_AUTOTUNE);
    for (auto& g : by_wm) {
        std::vector<HostInst> part;
        std::vector<drift_pt_result*> prs;
        int block = 0;
        if (by_wm.size() == 1 && big.empty() && !tune) {
            for (const HostInst& H : hs) block = std::max(block, H.block_req);
            const int rc = run_lds(hs, rs, block);
            if (rc) return rc;
            break;
        }
        part_of(g.second, part, prs, block);
        if (tune && block == 0) {
            BatchShape sh;
            const int rc = tune_batch(DRIFT_PT_ENGINE_LDS, part, sh);
            if (rc) return rc;
            block = sh.block;
        }
        const int rc = run_lds(part, prs, block);
        if (rc) return rc;
    }
    if (big.empty()) return 0;
    // the split engine when its workgroups (groups, times the parts asked for) fill the chip; the
    // fits are checked per instance
    std::vector<HostInst> cand;
    std::vector<size_t> cand_idx, rest;
    for (size_t i : big) {
        if (split_fits(hs[i])) {
            cand.push_back(hs[i]);
            cand_idx.push_back(i);
        } else {
            rest.push_back(i);
        }
    }
    if (!cand.empty()) {
        const std::vector<int> rgs = split_rgs(cand);
        int64_t groups = 0;
        for (size_t j = 0; j < cand.size(); ++j)
            groups += (int64_t)((cand[j].R + rgs[j] - 1) / rgs[j]) * std::max(1, cand[j].np_req);
        if (groups >= device_cus()) {
            for (size_t j = 0; j < cand.size(); ++j) split_wm[cand[j].wm].push_back(cand_idx[j]);
        } else {
            rest.insert(rest.end(), cand_idx.begin(), cand_idx.end());
        }
    }
    for (auto& g : split_wm) {
        std::vector<HostInst> part;
        std::vector<drift_pt_result*> prs;
        int block = 0;
        part_of(g.second, part, prs, block);
        bool given = block > 0;
        for (const HostInst& H : part) given = given || H.rg_req > 0 || H.np_req > 0;
        if (tune && !given) {
            BatchShape sh;
            const int rc = tune_batch(DRIFT_PT_ENGINE_SPLIT, part, sh);
            if (rc) return rc;
            block = sh.block;
            for (HostInst& H : part) {
                if (sh.rg_cap > 0) H.rg_req = std::max(1, std::min(sh.rg_cap, split_rg_max(H, sh.cl)));
                if (sh.cl) H.flags |= DRIFT_PT_COUPLINGS_LDS;
            }
        }
        const int rc = run_split(part, prs, block);
        if (rc) return rc;
    }
    if (rest.empty()) return 0;
    std::sort(rest.begin(), rest.end());
    std::vector<const HostInst*> gh;
    std::vector<drift_pt_result*> grs;
    for (size_t i : rest) {
        gh.push_back(&hs[i]);
        grs.push_back(rs[i]);
    }
    return run_global_set(gh, grs, 4, nullptr);
}

// The allocation-free entry point's device region.  Every device byte drift_pt_run_batch_noalloc
// uses -- the packed instance arrays, the instance descriptors, the work queue, the statistics
// blocks and the other outputs -- is a slice of it.  The module holds one copy per device from
// the time the library's code object is loaded on that device until the process ends.  Host code
// reaches it only through the per-device address hipGetSymbolAddress returns (noalloc_region).
static __device__ __align__(256) uint8_t scratch[DRIFT_PT_NOALLOC_BYTES];

struct NoallocSlot {
    std::mutex mu;            // one call at a time per device: calls share the region
    uint8_t* ptr = nullptr;   // device address of this device's copy of scratch
};
NoallocSlot g_noalloc[64];

// The slot of the current device, its region's address looked up on first use; null (with the
// error set) when that fails.
NoallocSlot* noalloc_region() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        fail(DRIFT_PT_EHIP, "allocation-free batch: no current device");
        return nullptr;
    }
    NoallocSlot& S = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(S.mu);
    if (!S.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fail(DRIFT_PT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        S.ptr = (uint8_t*)p;
    }
    return &S;
}

}  // namespace

extern "C" {

int drift_pt_run_batch(const drift_pt_problem* problems, const drift_pt_params* params,
                       int32_t n_instances, drift_pt_result* results) {
    g_err.clear();
    std::vector<HostInst> hs;
    std::vector<drift_pt_result*> rs;
    int rc = prepare_batch(problems, params, n_instances, results, hs, rs, false);
    if (rc) return rc;
    return run_batch(hs, rs);
}

int drift_pt_run_batch_noalloc(const drift_pt_problem* problems, const drift_pt_params* params,
                               int32_t n_instances, drift_pt_result* results) {
    g_err.clear();
    std::vector<HostInst> hs;
    std::vector<drift_pt_result*> rs;
    int rc = prepare_batch(problems, params, n_instances, results, hs, rs, true);
    if (rc) return rc;
    for (const HostInst& H : hs)
        if (H.wm != hs[0].wm)
            return fail(DRIFT_PT_EINVAL, "allocation-free batch: all instances need one weight_type");
    NoallocSlot* S = noalloc_region();
    if (!S) return DRIFT_PT_EHIP;
    std::lock_guard<std::mutex> lk(S->mu);
    DevArena A;
    A.region = S->ptr;
    A.cap = DRIFT_PT_NOALLOC_BYTES;
    rc = run_lds(hs, rs, 0, &A);
    if (rc && A.need > A.cap)
        return fail(DRIFT_PT_ERANGE,
                    "allocation-free batch: the instances need at least " + std::to_string(A.need) +
                        " bytes of device memory, more than the " +
                        std::to_string((size_t)DRIFT_PT_NOALLOC_BYTES) + " bytes it holds");
    return rc;
}

}  // extern "C"
