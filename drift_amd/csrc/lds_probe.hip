// lds_probe.hip -- device probes, built into libdriftprobe.so; not part of the engine's C ABI
// (include/drift_pt.h).  The LDS read rate of this device, for the LDS roofline of the engines whose
// state lives in LDS (the batched LDS engine, the persistent slab engines), and the largest gap
// between the hardware exponential and the deterministic one over every float argument the float
// Metropolis decision can see, which the engine's fast decision relies on.
//
//   int drift_probe_lds(int kind, int iters, double* seconds, double* bytes, int64_t* instrs)
//     kind 0: 16-byte reads (ds_read_b128), kind 1: 1-byte reads (ds_read_u8).  Eight 256-thread
//     workgroups per CU each make `iters` reads per lane; seconds is the device time of the timed
//     launch (after one warm-up launch), bytes the bytes read, instrs the wave-level LDS read
//     instructions.  Returns 0, or a negative HIP error code.
//   int drift_probe_exp_margin(double* max_rel, float* worst_x, double* min_nonneg)
//     max over every float x in (-87, 0] of |p - pa| / pa, p = pt_expf_det(x), pa = pt_exp_hw(x),
//     and the x where it occurs; min_nonneg is the least pt_exp_hw(x) over every float x >= 0
//     (+inf included).  Returns 0, or a negative HIP error code.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string.h>

#include "pt_device.h"

namespace {

// Consecutive lanes read consecutive words, so a wave's read touches every bank once.  The words
// are folded into one XOR, stored only if it equals a value no lane's data produces, so the reads
// stay live and no global memory traffic is timed.
__global__ __launch_bounds__(256) void lds_read_b128(uint32_t* sink, int iters, uint32_t never) {
    __shared__ uint4 buf[4096];   // 64 KiB
    for (int i = threadIdx.x; i < 4096; i += 256)
        buf[i] = make_uint4((uint32_t)i, (uint32_t)i * 3u, (uint32_t)i * 5u, (uint32_t)i * 7u);
    __syncthreads();
    uint32_t acc = 0;
    const int idx = threadIdx.x;
#pragma unroll 8
    for (int i = 0; i < iters; ++i) {
        const uint4 v = buf[(idx + i * 256) & 4095];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == never) sink[blockIdx.x * 256 + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void lds_read_u8(uint32_t* sink, int iters, uint32_t never) {
    __shared__ uint8_t buf[65536];
    for (int i = threadIdx.x; i < 65536; i += 256) buf[i] = (uint8_t)(i * 7 + 3);
    __syncthreads();
    uint32_t acc = 0;
    const int idx = threadIdx.x;
#pragma unroll 8
    for (int i = 0; i < iters; ++i) acc += buf[(idx + i * 256) & 65535];
    if (acc == never) sink[blockIdx.x * 256 + threadIdx.x] = acc;
}

// Float x = -(bits i) for i in [0, 0x42AE0000): -0 up to the float just above -87.  Each lane keeps
// the largest (relative gap bits, i) pair of its strided share; the gap is non-negative, so its
// float bits order like the gap and the packed pair orders by gap, then by i.
constexpr uint32_t kExpArgs = 0x42AE0000u;

__global__ __launch_bounds__(256) void exp_margin(unsigned long long* best) {
    unsigned long long mine = 0;
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < kExpArgs; i += stride) {
        const float x = __uint_as_float(0x80000000u | i);
        const double p = (double)pt_expf_det(x);
        const double pa = (double)pt_dev::pt_exp_hw(x);
        const float rel = (float)(fabs(p - pa) / pa);
        const unsigned long long key = ((unsigned long long)__float_as_uint(rel) << 32) | i;
        mine = key > mine ? key : mine;
    }
    atomicMax(best, mine);
}

// Least pt_exp_hw(x) over x = +(bits i), i in [0, 0x7F800000]: +0 up to +inf.  The values are
// non-negative (or NaN, kept as the largest bits), so their float bits order like the values.
__global__ __launch_bounds__(256) void exp_nonneg_min(unsigned int* least) {
    unsigned int mine = 0xFFFFFFFFu;
    const uint32_t stride = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i <= 0x7F800000u; i += stride) {
        const float pa = pt_dev::pt_exp_hw(__uint_as_float(i));
        const unsigned int b = pa == pa ? __float_as_uint(pa) : 0xFFFFFFFFu;
        mine = b < mine ? b : mine;
    }
    atomicMin(least, mine);
}

}  // namespace

extern "C" int drift_probe_exp_margin(double* max_rel, float* worst_x, double* min_nonneg) {
    unsigned long long* best = nullptr;
    unsigned long long host = 0;
    unsigned int least_host = 0;
    hipError_t e = hipMalloc(&best, sizeof(*best) + sizeof(unsigned int));
    unsigned int* least = (unsigned int*)(best + 1);
    if (e == hipSuccess) e = hipMemset(best, 0, sizeof(*best));
    if (e == hipSuccess) e = hipMemset(least, 0xFF, sizeof(unsigned int));
    if (e == hipSuccess) {
        exp_margin<<<4096, 256>>>(best);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        exp_nonneg_min<<<4096, 256>>>(least);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&host, best, sizeof(host), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&least_host, least, sizeof(least_host), hipMemcpyDeviceToHost);
    if (best) (void)hipFree(best);
    if (e != hipSuccess) return -(int)e;
    float least_f;
    memcpy(&least_f, &least_host, sizeof(least_f));
    *min_nonneg = least_host == 0xFFFFFFFFu ? NAN : (double)least_f;
    const uint32_t relbits = (uint32_t)(host >> 32), i = (uint32_t)host;
    float rel;
    memcpy(&rel, &relbits, sizeof(rel));
    const uint32_t xbits = 0x80000000u | i;
    memcpy(worst_x, &xbits, sizeof(*worst_x));
    *max_rel = rel;
    return 0;
}

extern "C" int drift_probe_lds(int kind, int iters, double* seconds, double* bytes,
                               int64_t* instrs) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    const int grid = cus * 8;
    uint32_t* sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&sink, sizeof(uint32_t) * (size_t)grid * 256);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {   // warm-up, then the timed one
        if (pass == 1) e = hipEventRecord(e0, 0);
        if (e != hipSuccess) break;
        if (kind == 0) lds_read_b128<<<grid, 256>>>(sink, iters, 0xDEADBEEFu);
        else lds_read_u8<<<grid, 256>>>(sink, iters, 0xDEADBEEFu);
        e = hipGetLastError();
        if (e == hipSuccess && pass == 1) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (e != hipSuccess) return -(int)e;
    *seconds = ms * 1e-3;
    *bytes = (double)grid * 256.0 * (double)iters * (kind == 0 ? 16.0 : 1.0);
    *instrs = (int64_t)grid * 4 * iters;   // four waves per workgroup
    return 0;
}
