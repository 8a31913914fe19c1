// Exhaustive device checks of the specialised f32 forms in gps_devmath.h against the general correctly
// rounded operations they replace, over every argument the particle kernels pass them.  Built by
// __graft_entry__.build() into tests/libgps_mathcheck.so; tests/test_gpu_math.py calls it.
#include "../gps-pbvr-web_amd/csrc/gps_devmath.h"
#include "../gps-pbvr-web_amd/csrc/gps_fastcount.h"
#include <string.h>

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

// bf16_pack2 (the hardware conversion of the denoiser's activations) against the integer round-to-nearest-even of
// the bit pattern, for every float but NaN, in both halves of the pair
__global__ __launch_bounds__(BLOCK) void k_bf16(unsigned long long *bad) {
    const uint64_t n = 1ull << 32;
    for (uint64_t i = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i < n; i += (uint64_t)GRID * BLOCK) {
        const uint32_t u = (uint32_t)i, v = u ^ 0x80000000u;
        bool b = false;
        if ((u & 0x7FFFFFFFu) <= 0x7F800000u) {
            const uint32_t ru = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16, rv = (v + 0x7FFFu + ((v >> 16) & 1u)) >> 16;
            const uint32_t got = gpsm::bf16_pack2(__uint_as_float(u), __uint_as_float(v));
            b = (got & 0xFFFFu) != ru || (got >> 16) != rv;
        }
        count_bad(b, bad);
    }
}

// exp_mid(x) == exp(x) for every x in [-80, 0] (bits 0x80000000 .. bits(-80))
__global__ __launch_bounds__(BLOCK) void k_exp_mid(unsigned long long *bad) {
    const uint32_t lo = 0x80000000u, n = __float_as_uint(-80.0f) - lo + 1u;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i - threadIdx.x < n; i += GRID * BLOCK) {
        bool b = false;
        if (i < n) {
            const float x = __uint_as_float(lo + i);
            b = __float_as_uint(gpsm::exp_mid(x)) != __float_as_uint(gpsm::exp(x));
        }
        count_bad(b, bad);
    }
}

// div_by(a, b, rcp_refined(b)) against the correctly rounded a / b over pseudo-random pairs (a PCG hash of the
// index): b with any mantissa and sign and an exponent in [-60, 60] (extremes: only -60, -59, 59, 60), a with an
// exponent in [-149, 57] (extremes: -100, -99, 56, 57), a quarter of them below 2^-100, zeros and subnormals
// included.  Bad: different bits where |a| >= 2^-100; where |a| < 2^-100, more than an ulp apart or |q| >= 2^-39.
__device__ uint32_t mix32(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

__global__ __launch_bounds__(BLOCK) void k_div_by(uint32_t pass, int extremes, unsigned long long *bad) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x + pass * (uint32_t)(GRID * BLOCK);
    const uint32_t h1 = mix32(i * 2u + 1u), h2 = mix32(h1 ^ 0x9e3779b9u), h3 = mix32(h2 + 7u);
    int eb, ea;
    if (extremes) {
        const int eb4[4] = {-60, -59, 59, 60}, ea4[4] = {-100, -99, 56, 57};
        eb = eb4[h3 & 3u];
        ea = ea4[(h3 >> 2) & 3u];
    } else {
        eb = -60 + (int)(h3 % 121u);
        ea = ((h3 >> 8) & 3u) == 0u ? -149 + (int)((h3 >> 10) % 49u) : -100 + (int)((h3 >> 10) % 158u);
    }
    const float b = __uint_as_float((h1 & 0x807FFFFFu) | ((uint32_t)(eb + 127) << 23));
    float a;
    if (ea < -126) a = __uint_as_float((h2 & 0x807FFFFFu) >> (uint32_t)(-126 - ea));   // subnormal or zero
    else a = __uint_as_float((h2 & 0x807FFFFFu) | ((uint32_t)(ea + 127) << 23));
    if ((h3 >> 26) == 63u) a = 0.0f;
    const float q = gpsm::div_by(a, b, gpsm::rcp_refined(b)), e = a / b;
    bool b_ad;
    if (fabsf(a) >= 0x1p-100f) {
        b_ad = __float_as_uint(q) != __float_as_uint(e);
    } else {
        const uint32_t uq = __float_as_uint(q) & 0x7FFFFFFFu, ue = __float_as_uint(e) & 0x7FFFFFFFu;
        b_ad = (uq > ue ? uq - ue : ue - uq) > 1u || !(fabsf(q) < 0x1p-39f);
    }
    count_bad(b_ad, bad);
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

// ---------------------------------------------------------------- error maxima of the count-only forms
constexpr float PI_F = 3.14159265358979f;   // the particle kernels' pi

__device__ void max_into(double v, unsigned long long *out) {
    if (!(v == v)) v = 1e300;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

// which: 0 Box-Muller radius (|fast - exact| over the 2^24 clamped uniforms; out[1], out[2] = max radius
// exact / fast), 1 sin/cos of 2 pi u (|fast - exact| over the 2^24 uniforms), 2 / 3 e^(-r2/2) exact / fast
// against f64 over every f32 r2 >= 0 (out[0] relative error where e^(-r2/2) >= 2^-126, out[1] absolute error
// below), 4 / 5 extinction ratio exact / fast against f64 over every f32 x in [0, 1 - 1e-6] (out[0] relative
// error, out[1] relative error times x),
// 6 v_rcp_f32 against f64 over every f32 z whose reciprocal is normal (relative)
__global__ __launch_bounds__(BLOCK) void k_measure(int which, unsigned long long *out) {
    uint64_t lo = 0, n = 0;
    switch (which) {
    case 0: case 1: n = 1u << 24; break;
    case 2: case 3: n = 0x7f800001u; break;
    case 4: case 5: n = (uint64_t)__float_as_uint(1.0f - 1e-6f) + 1u; break;
    default: lo = 0x00800000u; n = 0x7e800000u - lo; break;
    }
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i < n; i += (uint64_t)GRID * BLOCK) {
        const uint32_t b = (uint32_t)(lo + i);
        if (which == 0) {
            float u = gpsm::fmax((float)b * (1.0f / 16777216.0f), 5.96e-8f);
            float re = gpsm::sqrt_normal(-2.0f * gpsm::log_pos(u)), rf = gpsf::bm_radius(u);
            m0 = fmax(m0, fabs((double)rf - (double)re));
            m1 = fmax(m1, (double)re);
            m2 = fmax(m2, (double)rf);
        } else if (which == 1) {
            float u = (float)b * (1.0f / 16777216.0f);
            float th = 2.0f * PI_F * u, se, ce;
            gpsm::sincos_fin(th, se, ce);
            m0 = fmax(m0, fmax(fabs((double)gpsf::sin_turn(u) - se), fabs((double)gpsf::cos_turn(u) - ce)));
        } else if (which == 2 || which == 3) {
            float r2 = __uint_as_float(b);
            float t = which == 2 ? gpsm::exp_nonpos(-0.5f * gpsm::fmax(r2, 0.0f)) : gpsf::exp_half_neg(r2);
            double e = exp(-0.5 * (double)r2);
            if (e >= 0x1p-126) m0 = fmax(m0, fabs((double)t - e) / e);
            else m1 = fmax(m1, fabs((double)t - e));
        } else if (which == 4 || which == 5) {
            float x = __uint_as_float(b);
            float h = which == 4 ? gpsm::extinction_over_linear(x) : gpsf::extinction_over_linear(x);
            double xd = (double)x, e = x == 0.0f ? 1.0 : -log1p(-xd) / xd, rel = fabs((double)h - e) / e;
            m0 = fmax(m0, rel);
            m1 = fmax(m1, rel * xd);
        } else {
            float z = __uint_as_float(b);
            m0 = fmax(m0, fabs((double)gpsf::rcp(z) * (double)z - 1.0));
        }
    }
    max_into(m0, out);
    max_into(m1, out + 1);
    max_into(m2, out + 2);
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

// The count-only path's error constants, in the order of k_measure's quantities: E_R, E_T, ETA_TE, TAU_E,
// ETA_TF, TAU_F, ETA_HE, ETA_HE_X, ETA_HF, ETA_HF_X, E_RCP, U_MAX, DELTA_N.
extern "C" int mathcheck_constants(double *out, int n) {
    const float c[] = {gpsf::E_R, gpsf::E_T, gpsf::ETA_TE, gpsf::TAU_E, gpsf::ETA_TF, gpsf::TAU_F, gpsf::ETA_HE,
                       gpsf::ETA_HE_X, gpsf::ETA_HF, gpsf::ETA_HF_X, gpsf::E_RCP, gpsf::U_MAX, gpsf::DELTA_N};
    const int m = (int)(sizeof(c) / sizeof(c[0]));
    for (int k = 0; k < n && k < m; k++) out[k] = (double)c[k];
    return m;
}

// Each returns 0 and the number of mismatching arguments in *bad, or -1 on a HIP error.
extern "C" int mathcheck_measure(int which, double *out3) {
    if (which < 0 || which > 6) return -1;
    unsigned long long *d = nullptr;
    if (hipMalloc(&d, 3 * sizeof(*d)) != hipSuccess) return -1;
    int rc = hipMemset(d, 0, 3 * sizeof(*d)) == hipSuccess ? 0 : -1;
    if (!rc) {
        hipLaunchKernelGGL(k_measure, dim3(GRID), dim3(BLOCK), 0, 0, which, d);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -1;
    }
    unsigned long long h[3] = {0, 0, 0};
    if (!rc && hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    (void)hipFree(d);
    for (int k = 0; k < 3; k++) memcpy(&out3[k], &h[k], sizeof(double));
    return rc;
}

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

extern "C" int mathcheck_bf16(unsigned long long *bad) {
    return run(bad, [](unsigned long long *d) { hipLaunchKernelGGL(k_bf16, dim3(GRID), dim3(BLOCK), 0, 0, d); });
}

extern "C" int mathcheck_exp_mid(unsigned long long *bad) {
    return run(bad, [](unsigned long long *d) { hipLaunchKernelGGL(k_exp_mid, dim3(GRID), dim3(BLOCK), 0, 0, d); });
}

extern "C" int mathcheck_div_by(int extremes, unsigned long long *bad) {
    // 2^32 pairs: 2048 passes of the GRID x BLOCK threads
    return run(bad, [=](unsigned long long *d) {
        for (uint32_t pass = 0; pass < 2048u; pass++)
            hipLaunchKernelGGL(k_div_by, dim3(GRID), dim3(BLOCK), 0, 0, pass, extremes, d);
    });
}

extern "C" int mathcheck_sincos(uint32_t hi_bits, unsigned long long *bad) {
    return run(bad, [=](unsigned long long *d) {
        hipLaunchKernelGGL(k_sincos, dim3(GRID), dim3(BLOCK), 0, 0, hi_bits, d);
    });
}
