// Exhaustive device checks of the specialised f32 forms in gps_devmath.h against the general correctly
// rounded operations they replace, over every argument the particle kernels pass them.  Built by
// __graft_entry__.build() into tests/libgps_mathcheck.so; tests/test_gpu_math.py calls it.
#include "../gps-pbvr-web_amd/csrc/gps_devmath.h"

namespace {

constexpr int BLOCK = 256;
constexpr int GRID = 8192;

__device__ void count_bad(bool bad, unsigned long long *out) {
    unsigned long long n = __popcll(__ballot(bad));
    if (n && (threadIdx.x & 63) == 0) atomicAdd(out, n);
}

// sqrt_normal(x) == sqrt(x) for every x with bits in [lo, hi]
__global__ __launch_bounds__(BLOCK) void k_sqrt(uint32_t lo, uint32_t hi, unsigned long long *bad) {
    const uint64_t n = (uint64_t)hi - lo + 1u;
    for (uint64_t i = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i - threadIdx.x < n; i += (uint64_t)GRID * BLOCK) {
        bool b = false;
        if (i < n) {
            float x = __uint_as_float(lo + (uint32_t)i);
            b = __float_as_uint(gpsm::sqrt_normal(x)) != __float_as_uint(__builtin_sqrtf(x));
        }
        count_bad(b, bad);
    }
}

// log_core's quotient f / (2 + f) for every mantissa (f = m - 1 after the sqrt(2) fold)
__global__ __launch_bounds__(BLOCK) void k_div_log(unsigned long long *bad) {
    const uint32_t n = 1u << 23;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i - threadIdx.x < n; i += GRID * BLOCK) {
        bool b = false;
        if (i < n) {
            float m = __uint_as_float(i | 0x3f800000u);
            if (m > gpsm::SQRT2) m = m * 0.5f;
            float f = m - 1.0f;
            b = __float_as_uint(gpsm::div_normal(f, 2.0f + f)) != __float_as_uint(f / (2.0f + f));
        }
        count_bad(b, bad);
    }
}

// the extinction-over-linear quotient -log(1 - min(x, 1 - 1e-6)) / x for every x in [1e-4, 1 - 1e-6]
__global__ __launch_bounds__(BLOCK) void k_div_eol(unsigned long long *bad) {
    const uint32_t lo = __float_as_uint(1e-4f), hi = __float_as_uint(1.0f - 1e-6f), n = hi - lo + 1u;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i - threadIdx.x < n; i += GRID * BLOCK) {
        bool b = false;
        if (i < n) {
            float x = __uint_as_float(lo + i);
            float num = -gpsm::log_pos(1.0f - gpsm::fmin(x, 1.0f - 1e-6f));
            b = __float_as_uint(gpsm::div_normal(num, x)) != __float_as_uint(num / x);
        }
        count_bad(b, bad);
    }
}

// sincos_fin's quadrant selects against a switch over the quadrant, for every x with bits in [0, hi]
__global__ __launch_bounds__(BLOCK) void k_sincos(uint32_t hi, unsigned long long *bad) {
    const uint64_t n = (uint64_t)hi + 1u;
    for (uint64_t i = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i - threadIdx.x < n; i += (uint64_t)GRID * BLOCK) {
        bool b = false;
        if (i < n) {
            float x = __uint_as_float((uint32_t)i);
            float s, c;
            gpsm::sincos_fin(x, s, c);
            int q;
            float r = gpsm::reduce(x, q);
            float ps = gpsm::sin_poly(r), pc = gpsm::cos_poly(r), es, ec;
            switch (q) {
            case 0: es = ps; ec = pc; break;
            case 1: es = pc; ec = -ps; break;
            case 2: es = -ps; ec = -pc; break;
            default: es = -pc; ec = ps; break;
            }
            b = __float_as_uint(s) != __float_as_uint(es) || __float_as_uint(c) != __float_as_uint(ec);
        }
        count_bad(b, bad);
    }
}

template <typename F>
int run(unsigned long long *host, F launch) {
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, sizeof(*d)) != hipSuccess) return -1;
    int rc = 0;
    if (hipMemset(d, 0, sizeof(*d)) != hipSuccess) rc = -1;
    if (!rc) {
        launch(d);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -1;
    }
    if (!rc && hipMemcpy(host, d, sizeof(*d), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    (void)hipFree(d);
    return rc;
}

}  // namespace

// Each returns 0 and the number of mismatching arguments in *bad, or -1 on a HIP error.
extern "C" int mathcheck_sqrt(uint32_t lo_bits, uint32_t hi_bits, unsigned long long *bad) {
    if (hi_bits < lo_bits) return -1;
    return run(bad, [=](unsigned long long *d) {
        hipLaunchKernelGGL(k_sqrt, dim3(GRID), dim3(BLOCK), 0, 0, lo_bits, hi_bits, d);
    });
}

extern "C" int mathcheck_div_log(unsigned long long *bad) {
    return run(bad, [](unsigned long long *d) { hipLaunchKernelGGL(k_div_log, dim3(GRID), dim3(BLOCK), 0, 0, d); });
}

extern "C" int mathcheck_div_eol(unsigned long long *bad) {
    return run(bad, [](unsigned long long *d) { hipLaunchKernelGGL(k_div_eol, dim3(GRID), dim3(BLOCK), 0, 0, d); });
}

extern "C" int mathcheck_sincos(uint32_t hi_bits, unsigned long long *bad) {
    return run(bad, [=](unsigned long long *d) {
        hipLaunchKernelGGL(k_sincos, dim3(GRID), dim3(BLOCK), 0, 0, hi_bits, d);
    });
}
