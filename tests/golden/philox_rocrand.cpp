// Prints Philox4x32-10 outputs from rocRAND's engine (rocrand_philox4x32_10.h), an implementation
// independent of this repository, as JSON lines {"ctr": [...], "key": [...], "out": [...]}.
// The engine's state maps seed -> key, subsequence -> counter words 2..3 and offset/4 -> counter
// words 0..1, so any counter is reachable by seed + subsequence + repeated discards.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>
#include <cstdint>

static void emit(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    rocrand_device::philox4x32_10_engine eng;
    const unsigned long long seed = (unsigned long long)k0 | ((unsigned long long)k1 << 32);
    const unsigned long long sub = (unsigned long long)c2 | ((unsigned long long)c3 << 32);
    eng.seed(seed, sub, 0);
    unsigned long long v = (unsigned long long)c0 | ((unsigned long long)c1 << 32);
    while (v > 0) {
        const unsigned long long step = v < 0x3FFFFFFFFFFFFFFFull ? v : 0x3FFFFFFFFFFFFFFFull;
        eng.discard(4 * step);
        v -= step;
    }
    const uint4 o = eng.next4();
    std::printf("{\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}\n",
                c0, c1, c2, c3, k0, k1, o.x, o.y, o.z, o.w);
}

int main() {
    emit(0, 0, 0, 0, 0, 0);
    emit(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    emit(0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u);
    uint32_t x = 0x9E3779B9u;   // xorshift32 for a spread of counters and keys
    auto nx = [&]() { x ^= x << 13; x ^= x >> 17; x ^= x << 5; return x; };
    for (int i = 0; i < 64; ++i) {
        uint32_t a = nx(), b = nx(), c = nx(), d = nx(), e = nx(), f = nx();
        emit(a, b, c, d, e, f);
    }
    // counters of the engine's own stream layout: (p>>2, slot, sweep, tag)
    for (uint32_t tag = 1; tag <= 3; ++tag)
        for (uint32_t s = 0; s < 3; ++s) emit(s * 977u, s, s * 12345u, tag, 0x0000002au, 0u);
    return 0;
}
