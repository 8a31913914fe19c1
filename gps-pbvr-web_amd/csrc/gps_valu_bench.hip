// Microbenchmark of VALU issue throughput on gfx950 (in libgps_atomicbench.so; measurement only, not part of the
// renderer's ABI).  The particle kernel is bound by VALU issue: its rate in lane-instructions per second is read
// against what this kernel reaches with the same instruction mixes at full occupancy.  Every thread runs 8
// independent dependency chains for `iters` iterations, so issue (not latency) limits it:
//   mode 0: f32 fused multiply-add (v_fma_f32), one per chain and iteration
//   mode 1: the 32-bit PCG hash of the particle streams (v_mul_lo_u32, v_mad_u32, shifts, xors)
//   mode 2: 32-bit integer add and xor with a variable shift (the hash without the multiplies)
//   modes 3-17: one instruction class per chain step, pinned with register-only inline asm (no memory operands):
//   3 v_add_u32, 4 v_xor_b32, 5 v_mul_lo_u32, 6 v_add_f32, 7 v_fma_f32 and v_add_u32 on alternate chains,
//   8 v_lshrrev_b32 by a constant, 9 v_lshrrev_b32 by a register, 10 v_bfe_u32, 11 v_alignbit_b32,
//   12 v_lshl_add_u32, 13 v_xad_u32, 14 v_mad_u32_u24, 15 v_mul_hi_u32, 16 v_lshl_or_b32,
//   17 v_xor_b32 and v_lshrrev_b32 by a constant on alternate chains
//   modes 18-20: the PCG hash of mode 1 written out instruction by instruction (register-only inline asm), two
//   chains in lockstep with the constants in scalar registers (18), four chains in lockstep (19), two chains
//   with the constants as literals (20)
//   modes 21-26: transcendentals, one per chain step: 21 v_exp_f32, 22 v_rcp_f32, 23 v_sqrt_f32, 24 v_sin_f32,
//   25 v_log_f32, 26 v_exp_f32 and v_fmac_f32 on alternate chains
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

// one PCG step of two chains in lockstep; c1, c2, k in scalar registers
__device__ inline void pcg_x2(uint32_t &a, uint32_t &b, uint32_t c1, uint32_t c2, uint32_t k) {
    uint32_t ta, tb;
    asm volatile(
        "v_mul_lo_u32 %0, %0, %4\n\t"
        "v_mul_lo_u32 %1, %1, %4\n\t"
        "v_add_u32_e32 %0, %6, %0\n\t"
        "v_add_u32_e32 %1, %6, %1\n\t"
        "v_lshrrev_b32_e32 %2, 28, %0\n\t"
        "v_lshrrev_b32_e32 %3, 28, %1\n\t"
        "v_add_u32_e32 %2, 4, %2\n\t"
        "v_add_u32_e32 %3, 4, %3\n\t"
        "v_lshrrev_b32_e32 %2, %2, %0\n\t"
        "v_lshrrev_b32_e32 %3, %3, %1\n\t"
        "v_xor_b32_e32 %0, %2, %0\n\t"
        "v_xor_b32_e32 %1, %3, %1\n\t"
        "v_mul_lo_u32 %0, %0, %5\n\t"
        "v_mul_lo_u32 %1, %1, %5\n\t"
        "v_lshrrev_b32_e32 %2, 22, %0\n\t"
        "v_lshrrev_b32_e32 %3, 22, %1\n\t"
        "v_xor_b32_e32 %0, %2, %0\n\t"
        "v_xor_b32_e32 %1, %3, %1"
        : "+v"(a), "+v"(b), "=&v"(ta), "=&v"(tb)
        : "s"(c1), "s"(c2), "s"(k));
}

// the same with the added constant as a literal (v_mul_lo_u32, VOP3, takes no literal)
__device__ inline void pcg_x2_lit(uint32_t &a, uint32_t &b, uint32_t c1, uint32_t c2) {
    uint32_t ta, tb;
    asm volatile(
        "v_mul_lo_u32 %0, %0, %4\n\t"
        "v_mul_lo_u32 %1, %1, %4\n\t"
        "v_add_u32_e32 %0, 0xac564b05, %0\n\t"
        "v_add_u32_e32 %1, 0xac564b05, %1\n\t"
        "v_lshrrev_b32_e32 %2, 28, %0\n\t"
        "v_lshrrev_b32_e32 %3, 28, %1\n\t"
        "v_add_u32_e32 %2, 4, %2\n\t"
        "v_add_u32_e32 %3, 4, %3\n\t"
        "v_lshrrev_b32_e32 %2, %2, %0\n\t"
        "v_lshrrev_b32_e32 %3, %3, %1\n\t"
        "v_xor_b32_e32 %0, %2, %0\n\t"
        "v_xor_b32_e32 %1, %3, %1\n\t"
        "v_mul_lo_u32 %0, %0, %5\n\t"
        "v_mul_lo_u32 %1, %1, %5\n\t"
        "v_lshrrev_b32_e32 %2, 22, %0\n\t"
        "v_lshrrev_b32_e32 %3, 22, %1\n\t"
        "v_xor_b32_e32 %0, %2, %0\n\t"
        "v_xor_b32_e32 %1, %3, %1"
        : "+v"(a), "+v"(b), "=&v"(ta), "=&v"(tb)
        : "s"(c1), "s"(c2));
}

// four chains in lockstep
__device__ inline void pcg_x4(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d, uint32_t c1, uint32_t c2,
                              uint32_t k) {
    uint32_t ta, tb, tc, td;
    asm volatile(
        "v_mul_lo_u32 %0, %0, %8\n\t"
        "v_mul_lo_u32 %1, %1, %8\n\t"
        "v_mul_lo_u32 %2, %2, %8\n\t"
        "v_mul_lo_u32 %3, %3, %8\n\t"
        "v_add_u32_e32 %0, %10, %0\n\t"
        "v_add_u32_e32 %1, %10, %1\n\t"
        "v_add_u32_e32 %2, %10, %2\n\t"
        "v_add_u32_e32 %3, %10, %3\n\t"
        "v_lshrrev_b32_e32 %4, 28, %0\n\t"
        "v_lshrrev_b32_e32 %5, 28, %1\n\t"
        "v_lshrrev_b32_e32 %6, 28, %2\n\t"
        "v_lshrrev_b32_e32 %7, 28, %3\n\t"
        "v_add_u32_e32 %4, 4, %4\n\t"
        "v_add_u32_e32 %5, 4, %5\n\t"
        "v_add_u32_e32 %6, 4, %6\n\t"
        "v_add_u32_e32 %7, 4, %7\n\t"
        "v_lshrrev_b32_e32 %4, %4, %0\n\t"
        "v_lshrrev_b32_e32 %5, %5, %1\n\t"
        "v_lshrrev_b32_e32 %6, %6, %2\n\t"
        "v_lshrrev_b32_e32 %7, %7, %3\n\t"
        "v_xor_b32_e32 %0, %4, %0\n\t"
        "v_xor_b32_e32 %1, %5, %1\n\t"
        "v_xor_b32_e32 %2, %6, %2\n\t"
        "v_xor_b32_e32 %3, %7, %3\n\t"
        "v_mul_lo_u32 %0, %0, %9\n\t"
        "v_mul_lo_u32 %1, %1, %9\n\t"
        "v_mul_lo_u32 %2, %2, %9\n\t"
        "v_mul_lo_u32 %3, %3, %9\n\t"
        "v_lshrrev_b32_e32 %4, 22, %0\n\t"
        "v_lshrrev_b32_e32 %5, 22, %1\n\t"
        "v_lshrrev_b32_e32 %6, 22, %2\n\t"
        "v_lshrrev_b32_e32 %7, 22, %3\n\t"
        "v_xor_b32_e32 %0, %4, %0\n\t"
        "v_xor_b32_e32 %1, %5, %1\n\t"
        "v_xor_b32_e32 %2, %6, %2\n\t"
        "v_xor_b32_e32 %3, %7, %3"
        : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "=&v"(ta), "=&v"(tb), "=&v"(tc), "=&v"(td)
        : "s"(c1), "s"(c2), "s"(k));
}

template <int MODE>
__global__ __launch_bounds__(256) void k_valu(uint32_t *__restrict__ out, uint32_t iters, uint32_t seed) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (MODE >= 21) {
        float x[8];
        const float c = 0.999f + (float)(seed & 1u) * 1e-7f;
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = 0.5f + (float)(tid & 255u) * 1e-4f + (float)j * 1e-3f;
        for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if constexpr (MODE == 21) asm volatile("v_exp_f32 %0, %0" : "+v"(x[j]));
                else if constexpr (MODE == 22) asm volatile("v_rcp_f32 %0, %0" : "+v"(x[j]));
                else if constexpr (MODE == 23) asm volatile("v_sqrt_f32 %0, %0" : "+v"(x[j]));
                else if constexpr (MODE == 24) asm volatile("v_sin_f32 %0, %0" : "+v"(x[j]));
                else if constexpr (MODE == 25) asm volatile("v_log_f32 %0, %0" : "+v"(x[j]));
                else if (j & 1) asm volatile("v_fmac_f32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else asm volatile("v_exp_f32 %0, %0" : "+v"(x[j]));
            }
        }
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) s += x[j];
        out[tid] = __float_as_uint(s);
    } else if constexpr (MODE >= 18) {
        uint32_t x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = tid * 8u + (uint32_t)j + seed;
        const uint32_t c1 = 747796405u ^ (seed & 0u), c2 = 277803737u ^ (seed & 0u), k = 2891336453u ^ (seed & 0u);
        for (uint32_t it = 0; it < iters; it++) {
            if constexpr (MODE == 18) {
#pragma unroll
                for (int j = 0; j < 8; j += 2) pcg_x2(x[j], x[j + 1], c1, c2, k);
            } else if constexpr (MODE == 19) {
#pragma unroll
                for (int j = 0; j < 8; j += 4) pcg_x4(x[j], x[j + 1], x[j + 2], x[j + 3], c1, c2, k);
            } else {
#pragma unroll
                for (int j = 0; j < 8; j += 2) pcg_x2_lit(x[j], x[j + 1], c1, c2);
            }
        }
        uint32_t s = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) s ^= x[j];
        out[tid] = s;
    } else if constexpr (MODE == 0) {
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
                else if constexpr (MODE == 7) {
                    if (j & 1) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                    else asm volatile("v_fma_f32 %0, %0, %1, %1" : "+v"(x[j]) : "v"(c));
                } else if constexpr (MODE == 8) asm volatile("v_lshrrev_b32 %0, 13, %0" : "+v"(x[j]));
                else if constexpr (MODE == 9) asm volatile("v_lshrrev_b32 %0, %1, %0" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 10) asm volatile("v_bfe_u32 %0, %0, 3, 20" : "+v"(x[j]));
                else if constexpr (MODE == 11) asm volatile("v_alignbit_b32 %0, %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 12) asm volatile("v_lshl_add_u32 %0, %0, 3, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 13) asm volatile("v_xad_u32 %0, %0, %1, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 14) asm volatile("v_mad_u32_u24 %0, %0, %1, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 15) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                else if constexpr (MODE == 16) asm volatile("v_lshl_or_b32 %0, %0, 3, %1" : "+v"(x[j]) : "v"(c));
                else {
                    if (j & 1) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x[j]) : "v"(c));
                    else asm volatile("v_lshrrev_b32 %0, 13, %0" : "+v"(x[j]));
                }
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
// out_ms receives the mean launch time.  Returns 0, -1 on invalid arguments (mode 0..26, blocks 1..65536, iters
// 1..65536, reps 1..100), or -2 on a HIP error.
extern "C" int gps_valu_bench(int mode, uint32_t blocks, uint32_t iters, int reps, float *out_ms) {
    if (mode < 0 || mode > 26 || blocks < 1 || blocks > 65536 || iters < 1 || iters > 65536 || reps < 1 || reps > 100 ||
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
        else if (mode == 7) hipLaunchKernelGGL(k_valu<7>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 8) hipLaunchKernelGGL(k_valu<8>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 9) hipLaunchKernelGGL(k_valu<9>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 10) hipLaunchKernelGGL(k_valu<10>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 11) hipLaunchKernelGGL(k_valu<11>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 12) hipLaunchKernelGGL(k_valu<12>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 13) hipLaunchKernelGGL(k_valu<13>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 14) hipLaunchKernelGGL(k_valu<14>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 15) hipLaunchKernelGGL(k_valu<15>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 16) hipLaunchKernelGGL(k_valu<16>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 17) hipLaunchKernelGGL(k_valu<17>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 18) hipLaunchKernelGGL(k_valu<18>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 19) hipLaunchKernelGGL(k_valu<19>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 20) hipLaunchKernelGGL(k_valu<20>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 21) hipLaunchKernelGGL(k_valu<21>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 22) hipLaunchKernelGGL(k_valu<22>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 23) hipLaunchKernelGGL(k_valu<23>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 24) hipLaunchKernelGGL(k_valu<24>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else if (mode == 25) hipLaunchKernelGGL(k_valu<25>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
        else hipLaunchKernelGGL(k_valu<26>, dim3(blocks), dim3(256), 0, nullptr, buf, iters, (uint32_t)r);
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
