// Microbenchmark of VALU issue throughput on gfx950 (in libgps_atomicbench.so; measurement only, not part of the
// renderer's ABI).  The particle kernel is bound by VALU issue: its rate in lane-instructions per second is read
// against what this kernel reaches with the same instruction mixes at full occupancy.  Every thread runs 8
// independent dependency chains for `iters` iterations, so issue (not latency) limits it:
//   mode 0: f32 fused multiply-add (v_fma_f32), one per chain and iteration
//   mode 1: the 32-bit PCG hash of the particle streams (v_mul_lo_u32, v_mad_u32, shifts, xors)
//   mode 2: 32-bit integer add and xor with a variable shift (the hash without the multiplies)
//   modes 3-7: one instruction class per chain step, pinned with register-only inline asm (no memory operands):
//   3 v_add_u32, 4 v_xor_b32, 5 v_mul_lo_u32, 6 v_add_f32, 7 v_fma_f32 and v_add_u32 on alternate chains
// The lane-instruction count of a launch is read from the SQ_INSTS_VALU counter (x 64); SQ_ACTIVE_INST_VALU2
// counts the quad-cycles in which a SIMD issued two VALU instructions.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ inline uint32_t pcg_h(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_valu(uint32_t *__restrict__ out, uint32_t iters, uint32_t seed) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (MODE == 0) {
        float x[8];
        const float m = 0.999f + (float)(seed & 1u) * 1e-7f, c = 1e-3f;
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = (float)(tid + (uint32_t)j);
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = __builtin_fmaf(x[j], m, c);
        }
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) s += x[j];
        out[tid] = __float_as_uint(s);
    } else if constexpr (MODE >= 3) {
        uint32_t x[8];
        const uint32_t c = 0x9e3779b9u ^ seed;
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = tid * 8u + (uint32_t)j + seed;
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if constexpr (MODE == 3) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 4) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 5) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 6) asm volatile("v_add_f32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else if (j & 1) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else asm volatile("v_fma_f32 %0, %0, %1, %1" : "+v"(x[j]) : "v"(c));
            }
        }
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) s ^= x[j];
        out[tid] = s;
    } else {
        uint32_t x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = tid * 8u + (uint32_t)j + seed;
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if constexpr (MODE == 1) x[j] = pcg_h(x[j]);
                else x[j] = ((x[j] >> ((x[j] >> 28u) + 4u)) ^ x[j]) + 0x9e3779b9u;
            }
        }
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) s ^= x[j];
        out[tid] = s;
    }
}

}  // namespace

// Times `reps` launches of blocks x 256 threads, each running 8 chains of `iters` iterations of mode's operation;
// out_ms receives the mean launch time.  Returns 0, -1 on invalid arguments (mode 0..7, blocks 1..65536, iters
// 1..65536, reps 1..100), or -2 on a HIP error.
extern "C" int gps_valu_bench(int mode, uint32_t blocks, uint32_t iters, int reps, float *out_ms) {
    if (mode < 0 || mode > 7 || blocks < 1 || blocks > 65536 || iters < 1 || iters > 65536 || reps < 1 || reps > 100 ||
        !out_ms)
        return -1;
    uint32_t *buf = nullptr;
    if (hipMalloc(&buf, (size_t)blocks * 256 * 4) != hipSuccess) return -2;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int rc = 0;
    float total = 0.0f;
    for (int r = 0; r <= reps && rc == 0; r++) {   // launch 0 warms up
        (void)hipEventRecord(e0, nullptr);
        if (mode == 0) hipLaunchKernelGGL(k_valu<0>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 1) hipLaunchKernelGGL(k_valu<1>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 2) hipLaunchKernelGGL(k_valu<2>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 3) hipLaunchKernelGGL(k_valu<3>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 4) hipLaunchKernelGGL(k_valu<4>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 5) hipLaunchKernelGGL(k_valu<5>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 6) hipLaunchKernelGGL(k_valu<6>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else hipLaunchKernelGGL(k_valu<7>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) rc = -2;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) total += ms;
    }
    if (hipGetLastError() != hipSuccess) rc = -2;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    *out_ms = total / (float)reps;
    return rc;
}
