// f32 arithmetic of the particle kernels (gfx950 device code).
//
// The WGSL builtins exp/log/sin/cos that shaders/gps_splat.wgsl relies on are
// driver approximations with no fixed bits.  Here they are one fixed
// algorithm built only from IEEE-754 binary32 +, -, *, /, sqrt and fused
// multiply-add, which the device evaluates with correct rounding (compiled
// with -ffp-contract=off and -fhip-fp32-correctly-rounded-divide-sqrt).  A CPU evaluation of the same
// steps therefore reproduces every result bit for bit.
//
//   exp: Cody-Waite reduction x = k ln2 + r (ln2 in two parts, the high part
//        with 9 trailing zero bits so k*LN2_HI is exact), degree-7 Taylor
//        polynomial in Horner form, 2^k applied through the exponent field.
//   log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//        s = (m-1)/(m+1), odd series through s^9, plus e ln2 in two parts.
//   sin/cos: reduction by pi/2 with a three-part constant (exact products for
//        |k| < 2^11), Taylor polynomials of degree 9 / 10 on |r| <= pi/4.
//   Every Horner step and reduction step is one explicit fused multiply-add
//   (fmaf: a single correct rounding, the same on the CPU and on the GPU).
//
// Accuracy: <= 2 ulp against correctly rounded results over the argument
// ranges the pipeline uses (sin/cos arguments are 2 pi u, u in [0, 1)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GPS_DEV __device__ __forceinline__

namespace gpsm {

constexpr float LN2_HI = 0x1.62e4p-1f;
constexpr float LN2_LO = 0x1.7f7d1cp-20f;
constexpr float LOG2E = 0x1.715476p+0f;
constexpr float PIO2_1 = 0x1.92p+0f;
constexpr float PIO2_2 = 0x1.fb4p-12f;
constexpr float PIO2_3 = 0x1.4442d2p-24f;
constexpr float TWO_OVER_PI = 0x1.45f306p-1f;
constexpr float SQRT2 = 0x1.6a09e6p+0f;

GPS_DEV uint32_t bits(float x) { return __float_as_uint(x); }
GPS_DEV float fbits(uint32_t u) { return __uint_as_float(u); }
GPS_DEV float rint(float x) { return __builtin_rintf(x); }
GPS_DEV float sqrt(float x) { return __builtin_sqrtf(x); }
GPS_DEV float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
GPS_DEV float fmin(float a, float b) { return a < b ? a : b; }
GPS_DEV float fmax(float a, float b) { return a > b ? a : b; }

// a / b where a, b and a / b are finite and far from the ends of the normal range (exponents within
// [-100, 100], b != 0).  For such operands v_div_scale leaves both unscaled, v_div_fmas reduces to a plain
// fused multiply-add and v_div_fixup returns its input, so the steps below are the correctly rounded division
// (-fhip-fp32-correctly-rounded-divide-sqrt) without those three instructions: the same bits.
GPS_DEV float div_normal(float a, float b) {
    float r = __builtin_amdgcn_rcpf(b);
    float e = fma(-b, r, 1.0f);
    r = fma(e, r, r);
    float q = a * r;
    e = fma(-b, q, a);
    q = fma(e, r, q);
    e = fma(-b, q, a);
    return fma(e, r, q);
}

// div_normal's first steps, which depend on b alone: 1 / b refined once (for a divisor shared by many divisions)
GPS_DEV float rcp_refined(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    return fma(fma(-b, r, 1.0f), r, r);
}

// div_normal's remaining steps with r = rcp_refined(b): the same operations, so the same bits as div_normal(a, b);
// for b in [2^-60, 2^60] and |a| <= 2^57 it is the correctly rounded a / b wherever |a| >= 2^-100 (checked by the
// device math tests), and within an ulp below
GPS_DEV float div_by(float a, float b, float r) {
    float q = a * r;
    float e = fma(-b, q, a);
    q = fma(e, r, q);
    e = fma(-b, q, a);
    return fma(e, r, q);
}

// sqrt(x) for a positive normal x in [2^-96, 2^126]: v_sqrt_f32 is within one ulp, and the neighbour whose
// residual x - s' s changes sign is the correctly rounded root.  These are the correctly rounded sqrt's steps
// without its subnormal prescale and its zero/infinity select, which do not apply to such x: the same bits.
GPS_DEV float sqrt_normal(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    float dn = __uint_as_float(__float_as_uint(s) - 1u), up = __uint_as_float(__float_as_uint(s) + 1u);
    float t = fma(-dn, s, x) <= 0.0f ? dn : s;
    return fma(-up, s, x) > 0.0f ? up : t;
}

// exp(x) / 2^kf for kf = rint(x log2 e): the reduction and the polynomial
GPS_DEV float exp_reduced(float x, float kf) {
    float r = fma(-kf, LN2_HI, x);
    r = fma(-kf, LN2_LO, r);
    float p = 0x1.a01a02p-13f;
    p = fma(p, r, 0x1.6c16c2p-10f);
    p = fma(p, r, 0x1.111112p-7f);
    p = fma(p, r, 0x1.555556p-5f);
    p = fma(p, r, 0x1.555556p-3f);
    p = fma(p, r, 0.5f);
    p = fma(p, r, 1.0f);
    p = fma(p, r, 1.0f);
    return p;
}

// exp of x in [-103.97, 88.72] (the checked entry points below handle the rest)
GPS_DEV float exp_core(float x) {
    float kf = rint(x * LOG2E);
    int k = (int)kf;
    float p = exp_reduced(x, kf);
    if (k > 127) return p * fbits(0x7f000000u) * fbits((uint32_t)(k - 127 + 127) << 23);
    if (k < -126) return p * fbits((uint32_t)(k + 64 + 127) << 23) * 0x1p-64f;
    return p * fbits((uint32_t)(k + 127) << 23);
}

GPS_DEV float exp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return fbits(0x7f800000u);
    if (x < -103.972084045410156f) return 0.0f;
    return exp_core(x);
}

// exp of a non-NaN x <= 88.72 (the pipeline calls it with x <= 0): the same bits as exp() there
GPS_DEV float exp_nonpos(float x) {
    if (x < -103.972084045410156f) return 0.0f;
    return exp_core(x);
}

// exp of a finite x in [-80, 0]: exp_core's steps with 2^k applied by v_ldexp_f32.  The result p 2^k is a
// normal float there, so the scaling is exact, as the multiplication by the power of two is: the same bits as
// exp() (checked for every such x by the device math tests), without its range checks.
GPS_DEV float exp_mid(float x) {
    const float kf = rint(x * LOG2E);
    return __builtin_ldexpf(exp_reduced(x, kf), (int)kf);
}

// log of a positive, normal, finite x times 2^-e0 (e0 undoes a subnormal prescale)
GPS_DEV float log_core(float x, int e0) {
    uint32_t u = bits(x);
    int e = e0 + (int)(u >> 23) - 127;
    float m = fbits((u & 0x007fffffu) | 0x3f800000u);
    if (m > SQRT2) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = div_normal(f, 2.0f + f);   // |f| in {0} u [2^-24, 0.42], 2 + f in [1.7, 2.5]
    float z = s * s;
    float q = 0x1.c71c72p-4f;
    q = fma(q, z, 0x1.24924ap-3f);
    q = fma(q, z, 0x1.99999ap-3f);
    q = fma(q, z, 0x1.555556p-2f);
    float s2 = s + s;
    float lm = fma(s2, z * q, s2);
    float ef = (float)e;
    return fma(ef, LN2_HI, fma(ef, LN2_LO, lm));
}

GPS_DEV float log(float x) {
    if (x != x) return x;
    if (x < 0.0f) return fbits(0x7fc00000u);
    if (x == 0.0f) return fbits(0xff800000u);
    if (x == fbits(0x7f800000u)) return x;
    if (bits(x) < 0x00800000u) return log_core(x * 0x1p23f, -23);
    return log_core(x, 0);
}

// log of a positive, normal, finite x: the same bits as log() there
GPS_DEV float log_pos(float x) { return log_core(x, 0); }

// h(x) = -ln(1 - x) / x, the extinction-over-linear ratio of gps_splat.wgsl:417-460, for x <= 1 - 1e-6
// (both call sites clamp x there): the log's argument is in [1e-6, 1], the quotient in [1e-4, 14]
GPS_DEV float extinction_over_linear(float x) {
    if (x < 1e-4f) return 1.0f + 0.5f * x;
    return div_normal(-log_pos(1.0f - fmin(x, 1.0f - 1e-6f)), x);
}

GPS_DEV float sin_poly(float r) {
    float z = r * r;
    float p = 0x1.71de3ap-19f;
    p = fma(p, z, -0x1.a01a02p-13f);
    p = fma(p, z, 0x1.111112p-7f);
    p = fma(p, z, -0x1.555556p-3f);
    return fma(r, z * p, r);
}

GPS_DEV float cos_poly(float r) {
    float z = r * r;
    float p = -0x1.27e4fcp-22f;
    p = fma(p, z, 0x1.a01a02p-16f);
    p = fma(p, z, -0x1.6c16c2p-10f);
    p = fma(p, z, 0x1.555556p-5f);
    p = fma(p, z, -0.5f);
    return fma(z, p, 1.0f);
}

GPS_DEV float reduce(float x, int &q) {
    float kf = rint(x * TWO_OVER_PI);
    float r = fma(-kf, PIO2_1, x);
    r = fma(-kf, PIO2_2, r);
    r = fma(-kf, PIO2_3, r);
    q = ((int)kf) & 3;
    return r;
}

// sin and cos of a finite argument (the pipeline always needs both)
GPS_DEV void sincos_fin(float x, float &s, float &c) {
    int q;
    float r = reduce(x, q);
    float ps = sin_poly(r), pc = cos_poly(r);
    // quadrant q: (s, c) = (ps, pc), (pc, -ps), (-ps, -pc), (-pc, ps), as selects and sign flips
    const bool odd = (q & 1) != 0;
    float a = odd ? pc : ps, b = odd ? ps : pc;
    s = fbits(bits(a) ^ ((uint32_t)(q & 2) << 30));
    c = fbits(bits(b) ^ ((uint32_t)((q + 1) & 2) << 30));
}

GPS_DEV void sincos(float x, float &s, float &c) {
    if (x != x) {
        s = x;
        c = x;
        return;
    }
    sincos_fin(x, s, c);
}

GPS_DEV float sin(float x) { float s, c; sincos(x, s, c); return s; }
GPS_DEV float cos(float x) { float s, c; sincos(x, s, c); return c; }

// WGSL u32(f): truncate toward zero, saturate to [0, 2^32-1], NaN -> 0
GPS_DEV uint32_t f2u(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

GPS_DEV float clamp(float x, float lo, float hi) { return fmin(fmax(x, lo), hi); }
GPS_DEV float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Two floats to a packed bf16 pair (first in the low half), round to nearest even: one v_cvt_pk_bf16_f32 on gfx950.
// For every finite float it equals the integer rounding u + 0x7FFF + ((u >> 16) & 1) of the bit pattern (checked
// over all 2^32 patterns but NaN by the device math tests).  Used for the denoiser's forward activations.
GPS_DEV uint32_t bf16_pack2(float a, float b) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    const bf16x2_t v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}

}  // namespace gpsm
