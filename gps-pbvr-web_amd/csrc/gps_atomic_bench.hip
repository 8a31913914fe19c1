// Microbenchmark of 64-bit atomicMin throughput on gfx950 (libgps_atomicbench.so; measurement only, not part
// of the renderer's ABI).  The particle kernel's output is one packed atomicMin((depth key << 32) | splat id)
// per on-screen particle into a depth|id buffer of 8 B per subpixel, issued as write_min does it: a relaxed
// load, then the atomic only when the new value is smaller.  AMD publishes no L2 atomic rate, so this kernel
// measures it: every lane of every wave issues `per_thread` atomics of random 64-bit values into a buffer of
// 2^log2_n u64, at addresses drawn from a window of 2^log2_window entries around a random base per wave and
// iteration (window 1: all 64 lanes on one address; 64..4096: the footprint of one splat's particles;
// window = n: uniform over the buffer).
//   mode 0: __hip_atomic_fetch_min, result unused (global_atomic_min_x2 without return)
//   mode 1: write_min's form: relaxed load, then the atomic only when the value is smaller
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// 32-bit integer finaliser (xorshift-multiply rounds); any well-mixed hash serves here
__device__ inline uint32_t mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x21f0aaadu;
    x ^= x >> 15;
    x *= 0x735a2d97u;
    x ^= x >> 15;
    return x;
}

__global__ __launch_bounds__(256) void k_ab_fill(uint64_t *__restrict__ buf, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        buf[i] = ~0ull;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_ab_atomic_min(uint64_t *__restrict__ buf, uint32_t log2_n,
                                                        uint32_t log2_window, uint32_t per_thread, uint32_t seed) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nmask = (1ull << log2_n) - 1ull;
    const uint64_t wmask = (1ull << log2_window) - 1ull;
    for (uint32_t j = 0; j < per_thread; j++) {
        const uint32_t h = mix(wave * 0x9e3779b9u + j * 0x85ebca6bu + seed);
        const uint64_t base = (((uint64_t)h << 16) ^ (uint64_t)mix(h)) & nmask & ~wmask;
        const uint32_t r = mix(h ^ (lane * 0xc2b2ae35u + 1u));
        const uint64_t idx = base + ((uint64_t)mix(r) & wmask);   // < 2^log2_n: base and offset are disjoint bits
        const uint64_t v = ((uint64_t)r << 32) | (uint64_t)(wave * 64u + lane);
        uint64_t *d = buf + idx;
        if (MODE == 0) {
            __hip_atomic_fetch_min(d, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t seen = __hip_atomic_load(d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (v < seen) __hip_atomic_fetch_min(d, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace

// Times `reps` launches of blocks x 256 threads, each thread issuing per_thread atomics, on a buffer of
// 2^log2_n u64 refilled with all-ones before every launch (the fill is not timed).  out_ms receives the mean
// launch time.  Returns 0, or -1 on invalid arguments (log2_n in 10..30, log2_window <= log2_n, mode 0 or 1,
// blocks in 1..65536, per_thread in 1..4096, reps in 1..100), or -2 on a HIP error.
extern "C" int gps_atomic_bench(uint32_t log2_n, uint32_t log2_window, int mode, uint32_t blocks,
                                uint32_t per_thread, int reps, float *out_ms) {
    if (log2_n < 10u || log2_n > 30u || log2_window > log2_n || (mode != 0 && mode != 1) || blocks < 1u ||
        blocks > 65536u || per_thread < 1u || per_thread > 4096u || reps < 1 || reps > 100 || !out_ms)
        return -1;
    const uint64_t n = 1ull << log2_n;
    uint64_t *buf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float total = 0.0f;
    int rc = 0;
    if (hipMalloc(&buf, n * sizeof(uint64_t)) != hipSuccess) return -2;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -2;
    for (int it = 0; rc == 0 && it <= reps; it++) {   // iteration 0 warms up
        k_ab_fill<<<2048, 256>>>(buf, n);
        (void)hipEventRecord(e0, 0);
        if (mode == 0)
            k_ab_atomic_min<0><<<blocks, 256>>>(buf, log2_n, log2_window, per_thread, 0x1234567u * (it + 1));
        else
            k_ab_atomic_min<1><<<blocks, 256>>>(buf, log2_n, log2_window, per_thread, 0x1234567u * (it + 1));
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
            rc = -2;
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (it > 0) total += ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (rc == 0) *out_ms = total / (float)reps;
    return rc;
}
