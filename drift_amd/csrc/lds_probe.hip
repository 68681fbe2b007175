// lds_probe.hip -- the LDS read rate of this device, measured, for the LDS roofline of the engines
// whose state lives in LDS (the batched LDS engine, the persistent slab engines).  Built into
// libdriftprobe.so for the benchmark; not part of the engine's C ABI (include/drift_pt.h).
//
//   int drift_probe_lds(int kind, int iters, double* seconds, double* bytes, int64_t* instrs)
//     kind 0: 16-byte reads (ds_read_b128), kind 1: 1-byte reads (ds_read_u8).  Eight 256-thread
//     workgroups per CU each make `iters` reads per lane; seconds is the device time of the timed
//     launch (after one warm-up launch), bytes the bytes read, instrs the wave-level LDS read
//     instructions.  Returns 0, or a negative HIP error code.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

}  // namespace

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
