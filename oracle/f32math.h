/*
 * f32 elementary functions of the CPU oracle (test infrastructure only).
 *
 * The WGSL builtins exp/log/sin/cos (shaders/gps_splat.wgsl) have no bit-exact
 * definition: each WebGPU driver supplies its own approximation.  The HIP
 * kernels therefore evaluate these functions with one fixed algorithm built
 * only from IEEE-754 binary32 +, -, *, /, sqrt and fused multiply-add (all
 * correctly rounded on x86-64 and on gfx950 with -ffp-contract=off), so that
 * the oracle can reproduce every bit.  This file is the oracle's own statement of that
 * algorithm:
 *
 *   exp: Cody & Waite argument reduction x = k ln2 + r with ln2 split in two
 *        (W. J. Cody, W. Waite, "Software Manual for the Elementary
 *        Functions", 1980), degree-7 Taylor polynomial of e^r on |r| <= ln2/2,
 *        scaling by 2^k through the exponent field.
 *   log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s) with
 *        s = (m-1)/(m+1), odd series up to s^9 (J.-M. Muller, "Elementary
 *        Functions", ch. 6), plus e ln2 with the same two-part ln2.
 *   sin/cos: Cody & Waite reduction by pi/2 with a three-part constant,
 *        Taylor polynomials of degree 9 (sin) and 10 (cos) on |r| <= pi/4,
 *        quadrant selection.  Accurate for |x| < 2^14, which covers every
 *        argument the pipeline produces (2 pi u with u in [0,1)).
 *
 * Every Horner step and reduction step is one explicit fused multiply-add
 * (fmaf, correctly rounded on both the CPU and the GPU).
 *
 * Accuracy against libm is pinned by tests/test_oracle_mirror.py.
 * Compile with -ffp-contract=off (no FMA contraction) and without -ffast-math.
 */
#ifndef GPS_ORACLE_F32MATH_H
#define GPS_ORACLE_F32MATH_H

#include <stdint.h>
#include <math.h>
#include <string.h>

static inline uint32_t om_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float om_float(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

#define OM_LN2_HI 0x1.62e4p-1f          /* 0.693145751953125, 9 trailing zero bits */
#define OM_LN2_LO 0x1.7f7d1cp-20f
#define OM_LOG2E 0x1.715476p+0f
#define OM_PIO2_1 0x1.92p+0f            /* 1.5703125 */
#define OM_PIO2_2 0x1.fb4p-12f
#define OM_PIO2_3 0x1.4442d2p-24f
#define OM_TWO_OVER_PI 0x1.45f306p-1f
#define OM_SQRT2 0x1.6a09e6p+0f

static inline float om_rint(float x) { return __builtin_rintf(x); }
static inline float om_fma(float a, float b, float c) { return fmaf(a, b, c); }

static inline float om_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return om_float(0x7f800000u);
    if (x < -103.972084045410156f) return 0.0f;
    float kf = om_rint(x * OM_LOG2E);
    int k = (int)kf;
    float r = om_fma(-kf, OM_LN2_HI, x);
    r = om_fma(-kf, OM_LN2_LO, r);
    float p = 0x1.a01a02p-13f;                  /* 1/7! */
    p = om_fma(p, r, 0x1.6c16c2p-10f);          /* 1/6! */
    p = om_fma(p, r, 0x1.111112p-7f);           /* 1/5! */
    p = om_fma(p, r, 0x1.555556p-5f);           /* 1/4! */
    p = om_fma(p, r, 0x1.555556p-3f);           /* 1/3! */
    p = om_fma(p, r, 0.5f);
    p = om_fma(p, r, 1.0f);
    p = om_fma(p, r, 1.0f);
    if (k > 127) return p * om_float(0x7f000000u) * om_float((uint32_t)(k - 127 + 127) << 23);
    if (k < -126) return p * om_float((uint32_t)(k + 64 + 127) << 23) * 0x1p-64f;
    return p * om_float((uint32_t)(k + 127) << 23);
}

static inline float om_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return om_float(0x7fc00000u);
    if (x == 0.0f) return om_float(0xff800000u);
    if (x == om_float(0x7f800000u)) return x;
    int e = 0;
    uint32_t u = om_bits(x);
    if (u < 0x00800000u) {              /* subnormal: scale into the normal range */
        x = x * 0x1p23f;
        u = om_bits(x);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = om_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > OM_SQRT2) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float q = 0x1.c71c72p-4f;                   /* 1/9 */
    q = om_fma(q, z, 0x1.24924ap-3f);           /* 1/7 */
    q = om_fma(q, z, 0x1.99999ap-3f);           /* 1/5 */
    q = om_fma(q, z, 0x1.555556p-2f);           /* 1/3 */
    float s2 = s + s;
    float lm = om_fma(s2, z * q, s2);
    float ef = (float)e;
    return om_fma(ef, OM_LN2_HI, om_fma(ef, OM_LN2_LO, lm));
}

static inline float om_sin_poly(float r) {
    float z = r * r;
    float p = 0x1.71de3ap-19f;                  /* 1/9! */
    p = om_fma(p, z, -0x1.a01a02p-13f);         /* -1/7! */
    p = om_fma(p, z, 0x1.111112p-7f);           /* 1/5! */
    p = om_fma(p, z, -0x1.555556p-3f);          /* -1/3! */
    return om_fma(r, z * p, r);
}

static inline float om_cos_poly(float r) {
    float z = r * r;
    float p = -0x1.27e4fcp-22f;                 /* -1/10! */
    p = om_fma(p, z, 0x1.a01a02p-16f);          /* 1/8! */
    p = om_fma(p, z, -0x1.6c16c2p-10f);         /* -1/6! */
    p = om_fma(p, z, 0x1.555556p-5f);           /* 1/4! */
    p = om_fma(p, z, -0.5f);
    return om_fma(z, p, 1.0f);
}

/* quadrant k (mod 4) and reduced argument r of x = k pi/2 + r */
static inline float om_reduce(float x, int *quadrant) {
    float kf = om_rint(x * OM_TWO_OVER_PI);
    float r = om_fma(-kf, OM_PIO2_1, x);
    r = om_fma(-kf, OM_PIO2_2, r);
    r = om_fma(-kf, OM_PIO2_3, r);
    *quadrant = ((int)kf) & 3;
    return r;
}

static inline float om_sinf(float x) {
    if (x != x) return x;
    int q;
    float r = om_reduce(x, &q);
    switch (q) {
    case 0: return om_sin_poly(r);
    case 1: return om_cos_poly(r);
    case 2: return -om_sin_poly(r);
    default: return -om_cos_poly(r);
    }
}

static inline float om_cosf(float x) {
    if (x != x) return x;
    int q;
    float r = om_reduce(x, &q);
    switch (q) {
    case 0: return om_cos_poly(r);
    case 1: return -om_sin_poly(r);
    case 2: return -om_cos_poly(r);
    default: return om_sin_poly(r);
    }
}

/* WGSL u32(f): truncation toward zero, saturating at 0 and 2^32-1 (NaN -> 0) */
static inline uint32_t om_f2u(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

static inline float om_min(float a, float b) { return a < b ? a : b; }
static inline float om_max(float a, float b) { return a > b ? a : b; }
static inline float om_clamp(float x, float lo, float hi) { return om_min(om_max(x, lo), hi); }
static inline float om_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

#endif
