// Fast f32 forms for counting the particles of splats whose particles provably all land off-screen.
//
// Such a particle never writes a subpixel; its only effect is the generated-points statistic, i.e. whether
// it passes the near plane and (with the radial correction) the radial keep test.  The count-only path
// evaluates the particle with the hardware transcendentals below instead of the bit-exact gpsm functions,
// bounds the difference between the two evaluations of every quantity the decision depends on, and takes the
// fast result only when the decision cannot differ; every other particle is re-evaluated on the exact path.
//
// The error constants are maxima over every argument these functions receive, measured exhaustively on
// gfx950 against the bit-exact functions (radius, sin/cos: the 2^24 possible uniforms) or against f64
// (exponential, extinction ratio, reciprocal: every f32 argument), each rounded up and then doubled.
// tests/test_gpu_math.py measures them again and checks them against these constants.
#pragma once
#include "gps_devmath.h"

namespace gpsf {

constexpr float NEG_2LN2 = -0x1.62e430p+0f;         // -2 ln 2
constexpr float NEG_HALF_LOG2E = -0x1.715476p-1f;   // -log2(e) / 2
constexpr float LN2 = 0x1.62e430p-1f;

// sqrt(-2 ln u) for a clamped uniform u in [5.96e-8, 1)
GPS_DEV float bm_radius(float u) { return __builtin_amdgcn_sqrtf(NEG_2LN2 * __builtin_amdgcn_logf(u)); }

// sin(2 pi u) and cos(2 pi u) for a uniform u in [0, 1) (v_sin_f32 / v_cos_f32 take turns)
GPS_DEV float sin_turn(float u) { return __builtin_amdgcn_sinf(u); }
GPS_DEV float cos_turn(float u) { return __builtin_amdgcn_cosf(u); }

// e^(-r2 / 2) for r2 >= 0
GPS_DEV float exp_half_neg(float r2) { return __builtin_amdgcn_exp2f(NEG_HALF_LOG2E * r2); }

// -ln(1 - x) / x for x in [0, 1 - 1e-6], with the exact form's branch at 1e-4
GPS_DEV float extinction_over_linear(float x) {
    if (x < 1e-4f) return 1.0f + 0.5f * x;
    return -LN2 * __builtin_amdgcn_logf(1.0f - x) * __builtin_amdgcn_rcpf(x);
}

GPS_DEV float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

constexpr float EPS = 0x1p-24f;          // unit roundoff of f32 round-to-nearest

// Measured maxima (doubled):
constexpr float E_R = 1.0e-6f;           // |bm_radius - exact radius|, exact radius <= 5.7682 (measured 4.77e-7)
constexpr float E_T = 1.0e-6f;           // |sin_turn, cos_turn - exact sin, cos| (4.62e-7)
constexpr float ETA_TE = 1.6e-7f;        // exact e^(-r2/2): relative error where e^(-r2/2) >= 2^-126 (7.9e-8)
constexpr float TAU_E = 1.0e-44f;        //   absolute error below 2^-126 (1.2e-45)
constexpr float ETA_TF = 8.0e-6f;        // exp_half_neg: relative error (3.85e-6)
constexpr float TAU_F = 2.4e-38f;        //   absolute error below 2^-126 (1.18e-38)
constexpr float ETA_HE = 6.0e-4f;        // exact extinction ratio: relative error (2.98e-4) ...
constexpr float ETA_HE_X = 2.4e-7f;      //   ... and relative error times x (1.18e-7)
constexpr float ETA_HF = 6.0e-4f;        // fast extinction ratio: relative error (2.98e-4) ...
constexpr float ETA_HF_X = 5.3e-7f;      //   ... and relative error times x (2.63e-7)
constexpr float E_RCP = 2.0e-7f;         // v_rcp_f32 relative error (9.1e-8)

// Derived: every normal draw is at most U_MAX in magnitude on either path, and the two paths' draws differ
// by at most DELTA_N = R_MAX E_T + E_R (1 + E_T) + 2 EPS R_MAX (1 + E_T) = 7.46e-6 (the product's rounding
// on each path), rounded up.
constexpr float U_MAX = 5.78f;
constexpr float DELTA_N = 8.0e-6f;

}  // namespace gpsf
