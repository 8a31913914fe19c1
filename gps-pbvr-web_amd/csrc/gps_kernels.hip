// Particle-path kernels for gfx950.
//
// Reference pipeline per ensemble (crates/gps-render/src/gps.rs:884-921,
// shaders/gps_splat.wgsl): clear -> prepare -> multi-level scan -> finalize ->
// particle_depth -> particle_color -> resolve, one dispatch each.  Here:
//
//   k_view        once per render call: projection, per-splat particle law,
//                 3D Cholesky and SH colour for the call's camera (the same
//                 bits the reference recomputes for every ensemble)
//   k_count_scan  per batch of ensembles: Poisson counts fused with a
//                 single-pass exclusive scan (decoupled look-back, ticketed
//                 tiles), particle totals and the u32-overflow skip
//   k_particles   per batch: grid-stride kernel over every
//                 particle of every ensemble of the batch; one packed
//                 64-bit atomicMin((depth key << 32) | splat id) per particle
//                 replaces the reference's two regenerating passes
//   k_resolve     per batch: ensemble accumulation in ensemble order, with
//                 the depth|id buffers reset to empty in the same pass
//
// Every f32 expression keeps the operation order of the WGSL it restates;
// see gps_devmath.h for the elementary functions.
#include "gps_devmath.h"
#include "gps_fastcount.h"
#include "gps_internal.h"

namespace {

using gpsm::clamp;
using gpsm::f2u;
using gpsm::fmax;
using gpsm::fmin;
using gpsm::mix;

constexpr float PI_F = 3.14159265358979f;
constexpr float DILOG_ONE = 1.64493406684823f;

// ------------------------------------------------------------------ RNG
// gps_splat.wgsl:57-76
GPS_DEV uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

GPS_DEV uint32_t particle_seed(uint32_t splat, uint32_t k, uint32_t frame) {
    uint32_t s = pcg(splat * 747796405u + k + 1u);
    s = pcg(s ^ (frame * 2654435761u + 1u));
    return s ? s : 1u;
}

// The input of pcg for its output: each step is a bijection of u32 (the multiplications by their inverses modulo
// 2^32; the xorshift by r = (state >> 28) + 4 >= 4 leaves the top four bits, which give r back, and is undone by
// xoring the shifts by r, 2r, ...).
GPS_DEV uint32_t pcg_inv(uint32_t out) {
    const uint32_t w = out ^ (out >> 22u);
    const uint32_t x = w * 0xACB86D69u;   // 277803737^-1 mod 2^32
    const uint32_t r = (x >> 28u) + 4u;
    uint32_t s = x;
    for (uint32_t sh = r; sh < 32u; sh += r) s ^= x >> sh;
    return (s - 2891336453u) * 0xCC2BFE9Du;   // 747796405^-1 mod 2^32
}

// Rng::next's uniform for the PCG output c
GPS_DEV float unif24(uint32_t c) { return (float)(c >> 8u) * (1.0f / 16777216.0f); }

// particle_seed with the ensemble term frame * 2654435761 + 1 computed once per ensemble
GPS_DEV uint32_t particle_seed_mixed(uint32_t splat, uint32_t k, uint32_t frame_mix) {
    uint32_t s = pcg(splat * 747796405u + k + 1u);
    s = pcg(s ^ frame_mix);
    return s ? s : 1u;
}

struct Rng {
    uint32_t s;
    float spare;
    bool has_spare;
    GPS_DEV explicit Rng(uint32_t seed) : s(seed), spare(0.0f), has_spare(false) {}
    GPS_DEV float next() {
        s = pcg(s);
        return unif24(s);
    }
    // gps_splat.wgsl:402-414 (Box-Muller, sine half kept)
    GPS_DEV float normal() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        float u1 = fmax(next(), 5.96e-8f);
        float u2 = next();
        // u1 in [5.96e-8, 1 - 2^-24]: log is negative and normal, -2 log u1 in [1.19e-7, 33.3]
        float r = gpsm::sqrt_normal(-2.0f * gpsm::log_pos(u1));
        float th = 2.0f * PI_F * u2;
        float sn, cs;
        gpsm::sincos_fin(th, sn, cs);
        spare = r * sn;
        has_spare = true;
        return r * cs;
    }
};

// ------------------------------------------------------------------ Li2 / radius law
// gps_splat.wgsl:79-140
GPS_DEV float dilog_series(float x) {
    float term = x, sum = x;
    for (uint32_t k = 2u; k < 32u; k++) {
        term = term * x;
        sum = sum + term / (float)(k * k);
    }
    return sum;
}

GPS_DEV float dilog(float x) {
    if (x >= 1.0f) return DILOG_ONE;
    if (x <= 0.0f) return 0.0f;
    if (x <= 0.5f) return dilog_series(x);
    return DILOG_ONE - gpsm::log(x) * gpsm::log(1.0f - x) - dilog_series(1.0f - x);
}

GPS_DEV float inv_dilog(float tgt, float hi_in) {
    if (tgt <= 0.0f) return 0.0f;
    float lo = 0.0f, hi = hi_in, y = hi * 0.5f;
    for (uint32_t i = 0; i < 24u; i++) {
        float f = dilog(y) - tgt;
        if (f > 0.0f) hi = y; else lo = y;
        float dfdy = 1.0f;
        if (y > 1e-6f && y < 0.9999999f) dfdy = -gpsm::log(1.0f - y) / y;
        float ny = y - f / dfdy;
        if (!(ny > lo && ny < hi)) ny = 0.5f * (lo + hi);
        float dd = ny - y;
        if (dd < 0.0f) dd = -dd;
        if (dd <= 1e-6f * (1.0f + y)) return ny;
        y = ny;
    }
    return y;
}

GPS_DEV float radius_lut(const float *__restrict__ lut, float o, float u) {
    if (o <= 0.0f) return 0.0f;
    float v = -gpsm::log_pos(fmax(1.0f - u, 5.96e-8f));
    float fo = (1.0f - gpsm::sqrt(fmax(1.0f - fmin(o, 1.0f), 0.0f))) * 64.0f;
    float fv = fmin(gpsm::sqrt(v) * (float)(128.0 / 4.123105625617661), 128.0f);
    uint32_t i0 = f2u(fo);
    if (i0 > 63u) i0 = 63u;
    uint32_t j0 = f2u(fv);
    if (j0 > 127u) j0 = 127u;
    float a = fo - (float)i0, b = fv - (float)j0;
    uint32_t k = i0 * 129u + j0;
    float r0 = mix(lut[k], lut[k + 1u], b);
    float r1 = mix(lut[k + 129u], lut[k + 130u], b);
    return mix(r0, r1, a);
}

// out of line: the exact-radius option's series and Newton steps would otherwise be unrolled into every GPS
// particle loop (about 25k instructions), while the default table path is a handful
__device__ __noinline__ float corrected_radius(float o, float u) {
    if (o <= 0.0f) return 0.0f;
    float d = dilog(o);
    float tgt = (1.0f - u) * d;
    float y = fmin(inv_dilog(tgt, o), o);
    if (y <= 0.0f) return 40.0f;
    return gpsm::sqrt(fmax(-2.0f * gpsm::log(y / o), 0.0f));
}

// gps_splat.wgsl:142-160
GPS_DEV uint32_t poisson(Rng &r, float lambda) {
    if (lambda <= 0.0f) return 0u;
    if (lambda < 30.0f) {
        float l = gpsm::exp_nonpos(-lambda);
        uint32_t k = 0u;
        float p = 1.0f;
        for (;;) {
            k = k + 1u;
            p = p * r.next();
            if (p <= l) break;
        }
        return k - 1u;
    }
    float u1 = fmax(r.next(), 5.96e-8f);
    float u2 = r.next();
    float sn, cs;
    gpsm::sincos_fin(2.0f * PI_F * u2, sn, cs);
    float z = gpsm::sqrt(-2.0f * gpsm::log_pos(u1)) * cs;
    return f2u(fmax(gpsm::rint(lambda + gpsm::sqrt(lambda) * z), 0.0f));
}

// ------------------------------------------------------------------ projection
struct Proj {
    bool ok;
    float mx, my, a, b, c, depth;
};

GPS_DEV float dot3(const float *r, float x, float y, float z) { return r[0] * x + r[1] * y + r[2] * z; }

// gps_splat.wgsl:165-206
GPS_DEV Proj project(const GpsFrame &f, float4 g0, float4 g1, float4 g2) {
    Proj pr;
    pr.ok = false;
    pr.mx = pr.my = pr.a = pr.b = pr.c = pr.depth = 0.0f;
    float px = g0.x - f.pos[0], py = g0.y - f.pos[1], pz = g0.z - f.pos[2];
    float camx = dot3(f.rot[0], px, py, pz), camy = dot3(f.rot[1], px, py, pz), camz = dot3(f.rot[2], px, py, pz);
    if (camz <= fmax(f.near_z, 1e-4f)) return pr;
    const float lim = 5.2f;
    float tx = clamp(camx / camz, -lim, lim) * camz;
    float ty = clamp(camy / camz, -lim, lim) * camz;
    float fx = f.intr[0], fy = f.intr[1];
    float iz = 1.0f / camz;
    float sx = fx * iz, sxz = -fx * tx * iz * iz;
    float sy = fy * iz, syz = -fy * ty * iz * iz;
    float m0[3], m1[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        m0[j] = f.rot[0][j] * sx + f.rot[2][j] * sxz;
        m1[j] = f.rot[1][j] * sy + f.rot[2][j] * syz;
    }
    float ql = gpsm::sqrt(g2.x * g2.x + g2.y * g2.y + g2.z * g2.z + g2.w * g2.w);
    float w = g2.x / ql, x = g2.y / ql, y = g2.z / ql, z = g2.w / ql;
    float c0x = 1.0f - 2.0f * (y * y + z * z), c0y = 2.0f * (x * y + w * z), c0z = 2.0f * (x * z - w * y);
    float c1x = 2.0f * (x * y - w * z), c1y = 1.0f - 2.0f * (x * x + z * z), c1z = 2.0f * (y * z + w * x);
    float c2x = 2.0f * (x * z + w * y), c2y = 2.0f * (y * z - w * x), c2z = 1.0f - 2.0f * (x * x + y * y);
    float s2x = gpsm::exp(2.0f * g1.x), s2y = gpsm::exp(2.0f * g1.y), s2z = gpsm::exp(2.0f * g1.z);
    float ux = dot3(m0, c0x, c0y, c0z), uy = dot3(m0, c1x, c1y, c1z), uz = dot3(m0, c2x, c2y, c2z);
    float vx = dot3(m1, c0x, c0y, c0z), vy = dot3(m1, c1x, c1y, c1z), vz = dot3(m1, c2x, c2y, c2z);
    float a = s2x * (ux * ux) + s2y * (uy * uy) + s2z * (uz * uz) + f.pos[3];
    float b = s2x * (ux * vx) + s2y * (uy * vy) + s2z * (uz * vz);
    float c = s2x * (vx * vx) + s2y * (vy * vy) + s2z * (vz * vz) + f.pos[3];
    float det = a * c - b * b;
    if (!(det > 0.0f)) return pr;
    pr.ok = true;
    pr.mx = fx * camx / camz + f.intr[2];
    pr.my = fy * camy / camz + f.intr[3];
    pr.a = a;
    pr.b = b;
    pr.c = c;
    pr.depth = camz;
    return pr;
}

// gps_splat.wgsl:210-251 (evaluated once per splat per view)
// `sh` holds splat `id`'s rest coefficients at sh + id * 3 * sh_stride
GPS_DEV float4 splat_colour(const GpsFrame &f, uint32_t id, float4 g0, float4 g3, const float *__restrict__ sh) {
    float res[3] = {g3.x, g3.y, g3.z};
    uint32_t deg = f.sh_degree;
    if (deg == 0u) return make_float4(fmax(res[0], 0.0f), fmax(res[1], 0.0f), fmax(res[2], 0.0f), 0.0f);
    float px = g0.x - f.pos[0], py = g0.y - f.pos[1], pz = g0.z - f.pos[2];
    float len = gpsm::sqrt(px * px + py * py + pz * pz);
    float x = px / len, y = py / len, z = pz / len;
    uint32_t r = f.sh_stride;
    const size_t o = (size_t)id * 3u * r;
    float out[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        const float *S = sh + o + (uint32_t)ch * r;
        float v = res[ch];
        v = v - 0.4886025119029199f * y * S[0] + 0.4886025119029199f * z * S[1] - 0.4886025119029199f * x * S[2];
        if (deg >= 2u) {
            float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            v = v + 1.0925484305920792f * xy * S[3] - 1.0925484305920792f * yz * S[4] +
                0.31539156525252005f * (2.0f * zz - xx - yy) * S[5] - 1.0925484305920792f * xz * S[6] +
                0.5462742152960396f * (xx - yy) * S[7];
            if (deg >= 3u) {
                v = v - 0.5900435899266435f * y * (3.0f * xx - yy) * S[8] + 2.890611442640554f * xy * z * S[9] -
                    0.4570457994644658f * y * (4.0f * zz - xx - yy) * S[10] +
                    0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * S[11] -
                    0.4570457994644658f * x * (4.0f * zz - xx - yy) * S[12] +
                    1.445305721320277f * z * (xx - yy) * S[13] - 0.5900435899266435f * x * (xx - 3.0f * yy) * S[14];
            }
        }
        out[ch] = fmax(v, 0.0f);
    }
    return make_float4(out[0], out[1], out[2], 0.0f);
}

// gps_splat.wgsl:288-300
GPS_DEV bool offscreen(const GpsFrame &f, const Proj &pr) {
    float ex = 5.9f * gpsm::sqrt(pr.a), ey = 5.9f * gpsm::sqrt(pr.c);
    return pr.mx + ex < 0.0f || pr.mx - ex >= (float)f.W || pr.my + ey < 0.0f || pr.my - ey >= (float)f.H;
}

GPS_DEV float expected_count(const GpsFrame &f, const Proj &pr, float o) {
    uint32_t spp = f.spp_side * f.spp_side;
    float det = pr.a * pr.c - pr.b * pr.b;
    return f.density * (float)spp * 2.0f * PI_F * gpsm::sqrt(det) * dilog(o);
}

// gps_splat.wgsl:417-460
GPS_DEV float extinction_term(float o) {
    float c = fmin(o, 1.0f - 1e-6f);
    if (c < 1e-3f) return c * (1.0f + c * (0.5f + c * 0.33333334f));
    return -gpsm::log(1.0f - c);
}

GPS_DEV float pixel_density_scale(const GpsFrame &f, float depth) {
    if (!(depth > fmax(f.near_z, 0.0f))) return 0.0f;
    float ratio = f.ref_len / depth;
    return fmin(ratio * ratio * f.intr[0] * f.intr[1], 1.0e12f);
}


GPS_DEV float pbvr_lambda(const GpsFrame &f, const Proj &pr, float o) {
    float g = extinction_term(o);
    if (f.method == M_PROPORTIONAL) g = fmin(o, 1.0f - 1e-6f);
    float scale = f.base_k;
    if (f.method != M_VIEW_CONDITIONED) {
        if (f.calib == 1u) {
            scale = scale * pixel_density_scale(f, f.object_depth);
        } else if (f.calib == 2u) {
            scale = scale * pixel_density_scale(f, pr.depth);
        } else if (f.calib == 3u) {
            float spp = (float)(f.spp_side * f.spp_side);
            scale = spp * 2.0f * PI_F * gpsm::sqrt(pr.a * pr.c - pr.b * pr.b);
        }
    }
    float n = f.density * g * scale;
    if (!(n > 0.0f) || n > 3.0e38f) return 0.0f;
    return n;
}

// gps_splat.wgsl:464-485
GPS_DEV void chol_of(float4 g1, float4 g2, float4 &ca, float4 &cb) {
    float ql = gpsm::sqrt(g2.x * g2.x + g2.y * g2.y + g2.z * g2.z + g2.w * g2.w);
    float w = g2.x / ql, x = g2.y / ql, y = g2.z / ql, z = g2.w / ql;
    float c0[3] = {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y)};
    float c1[3] = {2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x)};
    float c2[3] = {2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)};
    float sx = gpsm::exp(2.0f * g1.x), sy = gpsm::exp(2.0f * g1.y), sz = gpsm::exp(2.0f * g1.z);
    float a00 = sx * c0[0] * c0[0] + sy * c1[0] * c1[0] + sz * c2[0] * c2[0];
    float a10 = sx * c0[1] * c0[0] + sy * c1[1] * c1[0] + sz * c2[1] * c2[0];
    float a20 = sx * c0[2] * c0[0] + sy * c1[2] * c1[0] + sz * c2[2] * c2[0];
    float a11 = sx * c0[1] * c0[1] + sy * c1[1] * c1[1] + sz * c2[1] * c2[1];
    float a21 = sx * c0[2] * c0[1] + sy * c1[2] * c1[1] + sz * c2[2] * c2[1];
    float a22 = sx * c0[2] * c0[2] + sy * c1[2] * c1[2] + sz * c2[2] * c2[2];
    float l00 = gpsm::sqrt(fmax(a00, 1e-30f));
    float l10 = a10 / l00;
    float l20 = a20 / l00;
    float l11 = gpsm::sqrt(fmax(a11 - l10 * l10, 1e-30f));
    float l21 = (a21 - l20 * l10) / l11;
    float l22 = gpsm::sqrt(fmax(a22 - l20 * l20 - l21 * l21, 1e-30f));
    ca = make_float4(l00, l10, l11, 0.0f);
    cb = make_float4(l20, l21, l22, 0.0f);
}

// ------------------------------------------------------------------ depth|id minimum
GPS_DEV void write_min(uint64_t *__restrict__ d, uint64_t v) {
    // the buffer only decreases, so a stale cached value >= the current one: skipping when it is already
    // <= v is exact and saves the L2 atomic for occluded particles
    uint64_t seen = __hip_atomic_load(d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (v < seen) __hip_atomic_fetch_min(d, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// ------------------------------------------------------------------ block reductions
template <typename T>
GPS_DEV T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Camera-space box of a PBVR splat's particles.  A particle's world position is centre + L u, with L the
// Cholesky factor of the 3-D covariance and u three normal draws of magnitude at most gpsf::U_MAX (Rng::normal
// clamps u1, so |u| <= sqrt(-2 ln 5.96e-8) = 5.77), so its real camera-space offset along axis k is at most
// U_MAX sum_j |(R L)_kj|.  bnd[k] bounds, for one particle, the difference between emit_pbvr's f32 value of
// camera axis k and count_only's.  On either path, with row sums l_m = sum_j |L_mj| and p = w - pos: the sum
// L u carries at most 3 EPS U_MAX l_m, w = c + L u adds EPS |w_m| <= EPS (|c_m| + U_MAX l_m), p adds EPS |p_m|
// and the dot product 3 EPS sum_m |R_km| |p_m|, with |p_m| <= |c_m - pos_m| + U_MAX l_m: together
// EPS sum_m |R_km| (|c_m| + 4 |c_m - pos_m| + 8 U_MAX l_m) from the real value of its own draws, and the two
// paths' draws differ by at most DELTA_N each.  Both paths' values of every particle lie in
// [cc_k - ext_k, cc_k + ext_k].
struct CamBox {
    float cc[3], ext[3], bnd[3];
};

GPS_DEV CamBox cam_box(const GpsFrame &f, float4 g0, float4 ca, float4 cb) {
    using namespace gpsf;
    const float L[3][3] = {{ca.x, 0.0f, 0.0f}, {ca.y, ca.z, 0.0f}, {cb.x, cb.y, cb.z}};
    const float c[3] = {g0.x, g0.y, g0.z};
    const float px = g0.x - f.pos[0], py = g0.y - f.pos[1], pz = g0.z - f.pos[2];
    const float p[3] = {px, py, pz};
    float row[3];
    for (int m = 0; m < 3; m++)
        row[m] = fabsf(c[m]) + 4.0f * fabsf(p[m]) + 8.0f * U_MAX * (fabsf(L[m][0]) + fabsf(L[m][1]) + fabsf(L[m][2]));
    CamBox b;
    for (int k = 0; k < 3; k++) {
        b.cc[k] = dot3(f.rot[k], px, py, pz);
        float s = 0.0f, m = 0.0f;
        for (int j = 0; j < 3; j++)
            s += fabsf(f.rot[k][0] * L[0][j] + f.rot[k][1] * L[1][j] + f.rot[k][2] * L[2][j]);
        for (int q = 0; q < 3; q++) m += fabsf(f.rot[k][q]) * row[q];
        b.bnd[k] = 2.02f * EPS * m + 1.001f * DELTA_N * s;
        b.ext[k] = U_MAX * s * 1.0001f + 1e-6f * (fabsf(b.cc[k]) + 1.0f) + 2.0f * b.bnd[k];
    }
    return b;
}

// No particle of the box can land in the image: over the part of the box in front of the near plane the pixel
// coordinates take their extremes at its corners, the low-pass jitter adds at most U_MAX * jitter pixels, and a
// 2-pixel margin covers the rounding of the projection.
GPS_DEV bool box_offscreen(const GpsFrame &f, const CamBox &b) {
    const float zhi = b.cc[2] + b.ext[2];
    if (zhi <= f.near_z) return true;
    const float zlo = fmaxf(b.cc[2] - b.ext[2], fmaxf(f.near_z, 1e-6f));
    const float margin = 2.0f + gpsf::U_MAX * f.jitter * 1.0001f;
    float lo[2], hi[2];
    for (int a = 0; a < 2; a++) {
        const float x0 = b.cc[a] - b.ext[a], x1 = b.cc[a] + b.ext[a];
        const float q0 = x0 / zlo, q1 = x0 / zhi, q2 = x1 / zlo, q3 = x1 / zhi;
        lo[a] = f.intr[a] * fminf(fminf(q0, q1), fminf(q2, q3)) + f.intr[2 + a] - margin;
        hi[a] = f.intr[a] * fmaxf(fmaxf(q0, q1), fmaxf(q2, q3)) + f.intr[2 + a] + margin;
    }
    return hi[0] < 0.0f || lo[0] >= (float)f.W || hi[1] < 0.0f || lo[1] >= (float)f.H;
}

// ================================================================== k_view
// The block's splats and SH rest coefficients are contiguous; they are staged in LDS with 16-byte loads
// across the block (a per-thread gather of 180-byte records would re-fetch cache lines), and each thread
// then reads its own record from LDS.
// The view kernel's body for block `blk` of one view (k_view: one launch per view; k_view_fused: every item of a
// batch in one launch).
GPS_DEV void view_body(const GpsFrame &f, const float4 *__restrict__ splats, const float *__restrict__ sh,
                       float4 *__restrict__ view, uint2 *__restrict__ cinfo, float4 *__restrict__ colour,
                       uint32_t *__restrict__ perm, uint32_t blk) {
    // each thread reads its splat and its SH rest coefficients itself: no LDS staging, so the kernel's occupancy
    // is set by its registers alone and the strided loads are served from the caches
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const bool live = i < f.N;
    bool marked = false;
    if (live) {
        const float4 *gs = splats + 4u * (size_t)i;
        float4 g0 = gs[0], g1 = gs[1], g2 = gs[2];
        float4 g3 = gs[3];
        colour[i] = splat_colour(f, i, g0, g3, sh);
        uint2 ci = make_uint2(0u, KIND_NONE);
        Proj pr = project(f, g0, g1, g2);
        float o = clamp(g0.w, 0.0f, 1.0f);
        if (pr.ok && o > 0.0f) {
            float l00 = gpsm::sqrt(fmax(pr.a, 1e-30f));
            float l10 = pr.b / l00;
            float l11 = gpsm::sqrt(fmax(pr.c - l10 * l10, 1e-30f));
            float4 v0 = make_float4(pr.mx, pr.my, pr.depth, o);
            float4 v1 = make_float4(l00, l10, l11, 1.0f);
            float4 v2 = make_float4(0, 0, 0, 0), v3 = v2, v4 = v2, v5 = v2;
            if (f.method == M_GPS) {
                // prepare_gps, gps_splat.wgsl:535-550
                if (!offscreen(f, pr)) {
                    float en = expected_count(f, pr, o);
                    if (en > 0.0f) ci = make_uint2(gpsm::bits(en), KIND_POISSON);
                }
            } else if (f.method == M_OCCUPANCY) {
                // prepare_pbvr occupancy branch, gps_splat.wgsl:553-581
                float density = fmax(f.density, 1e-8f);
                float amin = fmax(1.0f - gpsm::exp(gpsm::log(1.0f - 1.0f / 255.0f) / density), 1e-8f);
                if (o > amin) {
                    float radius = gpsm::sqrt(2.0f * gpsm::log(o / amin));
                    float ex = radius * gpsm::sqrt(pr.a), ey = radius * gpsm::sqrt(pr.c);
                    float Wf = (float)f.W, Hf = (float)f.H;
                    float lox = clamp(floorf(pr.mx - ex), 0.0f, Wf), loy = clamp(floorf(pr.my - ey), 0.0f, Hf);
                    float hix = clamp(ceilf(pr.mx + ex), 0.0f, Wf), hiy = clamp(ceilf(pr.my + ey), 0.0f, Hf);
                    float szx = fmax(hix - lox, 0.0f), szy = fmax(hiy - loy, 0.0f);
                    uint32_t count = f2u(szx) * f2u(szy) * f.spp_side * f.spp_side;
                    if (count != 0u) {
                        float invdet = 1.0f / fmax(pr.a * pr.c - pr.b * pr.b, 1e-30f);
                        v1 = make_float4(lox, loy, szx, szy);
                        float peak_tau = -gpsm::log(1.0f - fmin(o, 1.0f - 1e-6f));
                        bool sparse = density * peak_tau < 1.0f;
                        v2 = make_float4(pr.c * invdet, -pr.b * invdet, pr.a * invdet, peak_tau);
                        v3 = make_float4(0.0f, 0.0f, 0.0f, sparse ? 1.0f : 0.0f);
                        if (sparse) ci = make_uint2(gpsm::bits((float)count * density * peak_tau), KIND_POISSON);
                        else ci = make_uint2(count, KIND_FIXED);
                    }
                }
            } else {
                // prepare_pbvr, gps_splat.wgsl:582-611
                float lambda = pbvr_lambda(f, pr, o);
                if (lambda > 0.0f) {
                    float det = pr.a * pr.c - pr.b * pr.b;
                    float keep = 1.0f;
                    uint32_t kind = KIND_POISSON;
                    if (f.method == M_VIEW_CONDITIONED) {
                        float spp = (float)(f.spp_side * f.spp_side);
                        float target = f.density * spp * 2.0f * PI_F * gpsm::sqrt(det) * dilog(o);
                        if (lambda > 0.5f) keep = clamp(target / lambda, 0.0f, 1.0f); else keep = 0.0f;
                        if (target > lambda) kind |= KIND_UNDERCOVERED;
                    }
                    float idet = 1.0f / fmax(det, 1e-30f);
                    v1.w = keep;
                    v2 = make_float4(g0.x, g0.y, g0.z, pr.c * idet);
                    float oo = fmin(o, 1.0f - 1e-6f);
                    v3 = make_float4(-pr.b * idet, pr.a * idet, gpsm::extinction_over_linear(oo), 0.0f);
                    chol_of(g1, g2, v4, v5);
                    ci = make_uint2(gpsm::bits(lambda), kind);
                    if (f.flags & (FLAG_CULL | FLAG_FAST)) {
                        const CamBox bx = cam_box(f, g0, v4, v5);
                        if (box_offscreen(f, bx)) {
                            if (f.flags & FLAG_CULL) ci = make_uint2(0u, KIND_NONE);
                            else {   // count_only's marker and camera-space bounds
                                v1.x = bx.bnd[0];
                                v1.y = bx.bnd[1];
                                v1.z = bx.bnd[2];
                                v3.w = 1.0f;
                                // every particle in front of the near plane: the marker is minus a lower bound
                                // of the radial keep ratio clamp(h(y) / h(oo), 0, 1) over y in [0, oo], since the
                                // exact h >= (1 - ETA_HE) h(0) = 1 - ETA_HE and the division rounds correctly
                                if (bx.cc[2] - bx.ext[2] > f.near_z + 1e-6f * (fabsf(bx.cc[2]) + 1.0f))
                                    v3.w = -fmin((1.0f - gpsf::ETA_HE) / v3.z * (1.0f - 1e-6f), 1.0f);
                                marked = true;
                            }
                        }
                    }
                }
            }
            float4 *v = view + (size_t)VIEW_VEC4 * i;
            v[0] = v0;
            v[1] = v1;
            v[2] = v2;
            v[3] = v3;
            v[4] = v4;
            v[5] = v5;
        }
        cinfo[i] = ci;
    }
    if (perm) {   // marked splats from the front, the others from the back; one 64-bit atomic per block
        __shared__ uint32_t s_cnt[2][4], s_base[2];
        const uint64_t bm = __ballot(live && marked), bu = __ballot(live && !marked);
        const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
        if (lane == 0) {
            s_cnt[0][wave] = (uint32_t)__popcll(bm);
            s_cnt[1][wave] = (uint32_t)__popcll(bu);
        }
        __syncthreads();
        if (threadIdx.x == 0) {   // both counters in one word: marked in the low half, the others in the high
            uint32_t tm = 0, tu = 0;
            for (uint32_t w = 0; w < blockDim.x / 64u; w++) {
                tm += s_cnt[0][w];
                tu += s_cnt[1][w];
            }
            unsigned long long *ctr = reinterpret_cast<unsigned long long *>(perm + perm_counter(f.N));
            const unsigned long long old = atomicAdd(ctr, ((unsigned long long)tu << 32) | tm);
            s_base[0] = (uint32_t)old;
            s_base[1] = (uint32_t)(old >> 32);
        }
        __syncthreads();
        uint32_t om = s_base[0], ou = s_base[1];
        for (uint32_t w = 0; w < wave; w++) {
            om += s_cnt[0][w];
            ou += s_cnt[1][w];
        }
        if (live) {
            if (marked) perm[om + (uint32_t)__popcll(bm & below)] = i;
            else perm[f.N - 1u - (ou + (uint32_t)__popcll(bu & below))] = i;
        }
    }
}

__global__ __launch_bounds__(256) void k_view(GpsFrame f, const float4 *__restrict__ splats,
                                              const float *__restrict__ sh, float4 *__restrict__ view,
                                              uint2 *__restrict__ cinfo, float4 *__restrict__ colour,
                                              uint32_t *__restrict__ perm) {
    view_body(f, splats, sh, view, cinfo, colour, perm, blockIdx.x);
}

// ================================================================== decoupled look-back
// status word: [63:38] epoch, [37:36] flag (1 aggregate, 2 inclusive prefix), [35:0] value: saturated at 2^36 - 1
// for the particle counts (a total above 2^32 - 1 marks the ensemble as skipped), or modulo 2^32 (WRAP32) for the
// plain u32 scan, whose output wraps like the reference's u32 arithmetic
constexpr uint64_t ST_VAL_MASK = (1ull << 36) - 1ull;
constexpr uint32_t ST_EPOCH_MASK = (1u << 26) - 1u;

GPS_DEV uint64_t sat_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return s > ST_VAL_MASK ? ST_VAL_MASK : s;
}

template <bool WRAP32>
GPS_DEV uint64_t lb_add(uint64_t a, uint64_t b) {
    if constexpr (WRAP32) return (a + b) & 0xFFFFFFFFull;
    else return sat_add(a, b);
}

// Called by one thread per tile: publishes the aggregate and returns the exclusive prefix of `tile`.
template <bool WRAP32 = false>
GPS_DEV uint64_t lookback(uint64_t *__restrict__ status, uint32_t tile, uint64_t agg, uint32_t epoch) {
    const uint64_t tag = (uint64_t)(epoch & ST_EPOCH_MASK) << 38;
    if constexpr (WRAP32) agg &= 0xFFFFFFFFull;
    else agg = agg > ST_VAL_MASK ? ST_VAL_MASK : agg;
    if (tile == 0u) {
        __hip_atomic_store(&status[0], tag | (2ull << 36) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0ull;
    }
    __hip_atomic_store(&status[tile], tag | (1ull << 36) | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t excl = 0ull;
    int32_t t = (int32_t)tile - 1;
    for (;;) {
        uint64_t w;
        for (;;) {
            w = __hip_atomic_load(&status[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(w >> 38) == (epoch & ST_EPOCH_MASK) && ((w >> 36) & 3ull) != 0ull) break;
            __builtin_amdgcn_s_sleep(1);
        }
        excl = lb_add<WRAP32>(excl, w & ST_VAL_MASK);
        if (((w >> 36) & 3ull) == 2ull) break;
        t--;
    }
    __hip_atomic_store(&status[tile], tag | (2ull << 36) | lb_add<WRAP32>(excl, agg), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

// Block-wide exclusive scan of one u64 per thread (256 threads = 4 waves); returns the block total.
GPS_DEV uint64_t block_exclusive_scan(uint64_t v, uint64_t &excl, uint64_t *s_wave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint64_t wave_base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; w++) {
        uint64_t s = s_wave[w];
        if (w < wave) wave_base += s;
        total += s;
    }
    excl = wave_base + inc - v;
    return total;
}

// ================================================================== k_count_scan
// The body for one tile of ensemble e (k_count_scan: grid = (tiles, ensembles of the launch); k_count_scan_fused:
// every (item, ensemble) of a batch round in one launch).  Tiles are drawn by ticket per ensemble.
GPS_DEV void count_scan_body(const GpsFrame &f, const uint2 *__restrict__ cinfo, uint32_t first_seed,
                             uint32_t *__restrict__ offsets, const BatchState &bs, uint32_t tiles, uint32_t epoch,
                             uint64_t *__restrict__ stats, const uint32_t *__restrict__ perm, uint32_t e) {
    const uint32_t seed = first_seed + e;
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_wave[SCAN_BLOCK / 64];
    __shared__ uint64_t s_prefix;
    __shared__ uint64_t s_stats[6][SCAN_BLOCK / 64];
    if (threadIdx.x == 0) s_tile = atomicAdd(&bs.tickets[e], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t N = f.N;
    const uint32_t base = tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    const bool gps = f.method == M_GPS, occ = f.method == M_OCCUPANCY;
    uint32_t cnt[SCAN_ITEMS];
    uint64_t tsum = 0, st_pts = 0, st_trunc = 0, st_drop = 0, st_vis = 0, st_cand = 0, st_under = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint32_t i = base + (uint32_t)j;
        uint32_t c = 0;
        if (i < N) {
            const uint32_t sp = perm ? perm[i] : i;   // position i holds splat sp
            uint2 ci = cinfo[sp];
            uint32_t kind = ci.y & 3u;
            if (kind != KIND_NONE) {
                st_vis += 1;
                if (ci.y & KIND_UNDERCOVERED) st_under += 1;
                uint32_t n;
                if (kind == KIND_FIXED) {
                    n = ci.x;
                } else {
                    Rng r(particle_seed(sp, 0u, seed));
                    n = poisson(r, gpsm::fbits(ci.x));
                }
                if (gps) st_pts += n; else st_cand += n;
                if (!occ && n > f.cap) {
                    st_trunc += 1;
                    st_drop += n - f.cap;
                    n = f.cap;
                }
                c = n;
            }
        }
        cnt[j] = c;
        tsum += c;
    }
    uint64_t texcl;
    uint64_t agg = block_exclusive_scan(tsum, texcl, s_wave);
    if (threadIdx.x == 0) s_prefix = lookback(bs.status + (size_t)e * tiles, tile, agg, epoch);
    // statistics: one atomic per counter per tile
    {
        uint64_t v[6] = {st_pts, st_trunc, st_drop, st_vis, st_cand, st_under};
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            uint64_t s = wave_sum(v[k]);
            if (lane == 0) s_stats[k][wave] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        uint64_t s = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; w++) s += s_stats[threadIdx.x][w];
        const int slot[6] = {ST_POINTS, ST_TRUNC, ST_DROPPED, ST_VISIBLE, ST_CAND, ST_UNDER};
        if (s) atomicAdd((unsigned long long *)&stats[slot[threadIdx.x]], (unsigned long long)s);
    }
    const uint64_t prefix = s_prefix;
    uint32_t *out = offsets + (size_t)e * N;
    uint64_t run = prefix + texcl;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint32_t i = base + (uint32_t)j;
        if (i < N) out[i] = (uint32_t)run;
        run += cnt[j];
    }
    if (tile == tiles - 1u && threadIdx.x == 0) {
        uint64_t total = prefix + agg;
        bool over = total > 0xffffffffull;
        bs.totals[e] = over ? 0u : (uint32_t)total;
        bs.skip[e] = over ? 1u : 0u;
        if (over) atomicAdd((unsigned long long *)&stats[ST_SKIPPED], 1ull);
        bs.tickets[e] = 0u;   // every tile of this ensemble has drawn its ticket
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_count_scan(GpsFrame f, const uint2 *__restrict__ cinfo,
                                                           uint32_t first_seed, uint32_t *__restrict__ offsets,
                                                           BatchState bs, uint32_t tiles, uint32_t epoch,
                                                           uint64_t *__restrict__ stats,
                                                           const uint32_t *__restrict__ perm) {
    count_scan_body(f, cinfo, first_seed, offsets, bs, tiles, epoch, stats, perm, blockIdx.y);
}

// ================================================================== particles
GPS_DEV uint32_t owner_search(const uint32_t *__restrict__ offs, uint32_t lo, uint32_t hi, uint32_t p) {
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (offs[mid] <= p) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// owner_search over [lo, hi] (offs[lo] <= p) by one whole wave (every lane calls it with the same arguments):
// each step samples 64 evenly spaced positions of the remaining range, one per lane, and keeps the interval after
// the last sample <= p, so N = 560k takes 4 dependent loads instead of 20.  With gallop, when the owner is
// expected just after lo, the first steps sample the 64 positions from lo, then every 64th of the next 4096.
GPS_DEV uint32_t owner_search_wave(const uint32_t *__restrict__ offs, uint32_t lo, uint32_t hi, uint32_t p,
                                   bool gallop) {
    const uint32_t lane = threadIdx.x & 63u;
    if (gallop) {
        for (uint32_t s = 1u; s <= 64u; s *= 64u) {
            const uint32_t q = lo + lane * s;
            const uint64_t m = __ballot(q <= hi && offs[q] <= p);   // lane 0 (q = lo) is always set
            const uint32_t last = 63u - (uint32_t)__clzll(m);
            if (last < 63u || hi - lo <= 63u * s) {   // the owner is in [lo + last s, lo + last s + s - 1]
                lo += last * s;
                hi = min(hi, lo + s - 1u);
                break;
            }
            lo += 63u * s;
        }
    }
    while (hi - lo >= 64u) {   // offs[lo] <= p, and the owner is in [lo, hi]
        const uint32_t step = (hi - lo) / 64u + 1u;   // 64 samples cover hi - lo + 1 positions
        const uint32_t q = lo + lane * step;
        const uint64_t m = __ballot(q <= hi && offs[q] <= p);
        const uint32_t last = 63u - (uint32_t)__clzll(m);
        lo += last * step;
        hi = min(hi, lo + step - 1u);
    }
    const uint32_t q = lo + lane;
    const uint64_t m = __ballot(q <= hi && offs[q] <= p);
    return lo + 63u - (uint32_t)__clzll(m);
}

// A generated particle: the subpixel it writes (NO_HIT if none) and its ordered depth key.
constexpr uint32_t NO_HIT = 0xFFFFFFFFu;
struct Hit {
    uint32_t idx;
    uint32_t key;
    uint32_t counted;   // 1 if the particle counts as a generated point (PBVR / occupancy statistic)
};

// gps_splat.wgsl:259-282 (GPS)
GPS_DEV Hit emit_gps(const GpsFrame &f, Rng &r, float4 v0, float4 v1, const float *__restrict__ lut) {
    Hit h{NO_HIT, 0xFFFFFFFFu, 0u};
    float u = r.next();
    float rad = f.exact_radius == 0u ? radius_lut(lut, v0.w, u) : corrected_radius(v0.w, u);
    float th = 2.0f * PI_F * r.next();
    float sn, cs;
    gpsm::sincos_fin(th, sn, cs);
    float wx = rad * cs, wy = rad * sn;
    float pxf = v0.x + v1.x * wx, pyf = v0.y + (v1.y * wx + v1.z * wy);
    h.key = gpsm::bits(v0.z) | 0x80000000u;
    if (pxf < 0.0f || pxf >= f.wf || pyf < 0.0f || pyf >= f.hf) return h;
    const uint32_t side = f.spp_side;
    uint32_t px = f2u(pxf), py = f2u(pyf);
    uint32_t sx = f2u((pxf - (float)px) * f.sidef), sy = f2u((pyf - (float)py) * f.sidef);
    if (sx > side - 1u) sx = side - 1u;
    if (sy > side - 1u) sy = side - 1u;
    h.idx = (py * f.W + px) * (side * side) + sy * side + sx;
    return h;
}

// gps_splat.wgsl:490-532 (PBVR)
GPS_DEV Hit emit_pbvr(const GpsFrame &f, Rng &r, const float4 *__restrict__ v) {
    Hit h{NO_HIT, 0xFFFFFFFFu, 0u};
    float uz = r.normal();
    float uy = r.normal();
    float ux = r.normal();
    float4 a = v[0];
    if (f.method == M_VIEW_CONDITIONED) {
        if (r.next() >= v[1].w) return h;
    }
    float4 c = v[2], ca = v[4], cb = v[5];
    float wx = c.x + ca.x * ux;
    float wy = c.y + (ca.y * ux + ca.z * uy);
    float wz = c.z + (cb.x * ux + cb.y * uy + cb.z * uz);
    float px = wx - f.pos[0], py = wy - f.pos[1], pz = wz - f.pos[2];
    float camx = dot3(f.rot[0], px, py, pz), camy = dot3(f.rot[1], px, py, pz), camz = dot3(f.rot[2], px, py, pz);
    if (camz <= f.near_z) return h;
    float pixx = f.intr[0] * camx / camz + f.intr[2];
    float pixy = f.intr[1] * camy / camz + f.intr[3];
    if (f.flags & FLAG_JITTER) {
        float jy = r.normal();
        float jx = r.normal();
        pixx = pixx + f.jitter * jx;
        pixy = pixy + f.jitter * jy;
    }
    if (f.flags & FLAG_RADIAL) {
        float4 d = v[3];
        float dvx = pixx - a.x, dvy = pixy - a.y;
        float r2 = c.w * dvx * dvx + 2.0f * d.x * dvx * dvy + d.y * dvy * dvy;
        // radial_keep, gps_splat.wgsl:434-438, with h(min(o, 1-1e-6)) cached per splat in d.z
        float oo = fmin(a.w, 1.0f - 1e-6f);
        float y = oo * gpsm::exp_nonpos(-0.5f * fmax(r2, 0.0f));
        float keep = clamp(gpsm::extinction_over_linear(y) / d.z, 0.0f, 1.0f);
        if (r.next() >= keep) return h;
    }
    float depth_z = (f.flags & FLAG_CENTRE) ? a.z : camz;
    h.counted = 1u;
    h.key = gpsm::bits(depth_z) | 0x80000000u;
    if (pixx < 0.0f || pixx >= f.wf || pixy < 0.0f || pixy >= f.hf) return h;
    const uint32_t side = f.spp_side;
    uint32_t ipx = f2u(pixx), ipy = f2u(pixy);
    uint32_t sx = f2u((pixx - (float)ipx) * f.sidef), sy = f2u((pixy - (float)ipy) * f.sidef);
    if (sx > side - 1u) sx = side - 1u;
    if (sy > side - 1u) sy = side - 1u;
    h.idx = (ipy * f.W + ipx) * (side * side) + sy * side + sx;
    return h;
}

// gps_splat.wgsl:686-708 (screen occupancy)
GPS_DEV Hit emit_occupancy(const GpsFrame &f, Rng &r, uint32_t k, const float4 *__restrict__ v) {
    Hit h{NO_HIT, 0xFFFFFFFFu, 0u};
    const uint32_t side = f.spp_side, samples = side * side;
    float4 a = v[0], b = v[1], inv = v[2];
    bool sparse = v[3].w > 0.0f;
    uint32_t cell = k;
    if (sparse) {
        uint32_t a1 = f2u(r.next() * b.z * b.w * (float)samples);
        uint32_t a2 = f2u(b.z) * f2u(b.w) * samples - 1u;
        cell = a1 < a2 ? a1 : a2;
    }
    uint32_t pixel = cell / samples, sub = cell % samples;
    uint32_t bz = f2u(b.z);
    uint32_t xyx = f2u(b.x) + pixel % bz, xyy = f2u(b.y) + pixel / bz;
    float n1 = r.next();
    float n2 = r.next();
    float uvx = ((float)(sub % side) + n1) / (float)side;
    float uvy = ((float)(sub / side) + n2) / (float)side;
    float dx = (float)xyx + uvx - a.x, dy = (float)xyy + uvy - a.y;
    float r2 = fmax(inv.x * dx * dx + 2.0f * inv.y * dx * dy + inv.z * dy * dy, 0.0f);
    float alpha = fmin(a.w * gpsm::exp_nonpos(-0.5f * r2), 1.0f - 1e-6f);
    float coverage;
    if (sparse) coverage = -gpsm::log_pos(1.0f - alpha) / inv.w;   // 1 - alpha in [1e-6, 1]
    else coverage = 1.0f - gpsm::exp(f.density * gpsm::log_pos(1.0f - alpha));
    if (r.next() >= coverage) return h;
    h.idx = (xyy * f.W + xyx) * samples + sub;
    h.key = gpsm::bits(a.z) | 0x80000000u;
    h.counted = 1u;
    return h;
}

// Count-only evaluation (gps_fastcount.h) of a particle of a splat that the view kernel marked off-screen
// (v[3].w != 0; no particle of it can write a subpixel): 1 or 0 when emit_pbvr's generated-points decision is
// certain, -1 when it is not and the particle must go through emit_pbvr.  Its uniforms are emit_pbvr's, taken
// from the particle's PCG chain as the caller drew it: c1..c4 the two Box-Muller pairs of the three normals;
// then, after the view-conditioned draw (which the caller took: only particles it keeps come here), the jitter
// normals (the second pair's sine and a third pair, c5 c6) and the radial keep draw ck.  For a splat whose
// particles all pass the near plane (v[3].w < 0) the caller drew the whole chain and t0, t1, t2 are c5, c6, ck;
// otherwise t0 is the state they follow from, hashed here only when the particle passes the near plane.
// Each decision compares the count-only value with a bound on its
// difference from emit_pbvr's: camz against near_z with margin bnd_z (v[1].z, cam_box); with the radial
// correction, the keep ratio K = clamp(h(min(o e^(-r2/2), 1 - 1e-6)) / d.z, 0, 1) against the keep draw with
// margin dk.  dk collects the pixel difference (bx, by: camera bounds bnd_x, bnd_y, bnd_z through
// intr * cam / camz with camz >= camz - bnd_z on both paths, plus the roundings of the projection (v_rcp_f32 on
// this path), the jitter and dv), the r2 difference through the conic (10 EPS of its absolute terms for the
// roundings of both evaluations), its effect on e^(-r2/2) and on h (h' <= 1/(1-x)), and the relative errors of
// both evaluations of e^(-r2/2), h and the divisions.
GPS_DEV int count_only(const GpsFrame &f, const float4 *__restrict__ v, uint32_t c1, uint32_t c2, uint32_t c3,
                       uint32_t c4, uint32_t t0, uint32_t t1, uint32_t t2) {
    // Fused multiply-adds only remove roundings from this path's arithmetic, which every bound below counts
    // as separate operations; the measured gpsf functions are defined outside this scope and stay unfused.
#pragma clang fp contract(fast)
    using namespace gpsf;
    const float u1 = fmax(unif24(c1), 5.96e-8f), u2 = unif24(c2);
    const float u3 = fmax(unif24(c3), 5.96e-8f), u4 = unif24(c4);
    const float4 b = v[1];
    const float r1 = bm_radius(u1), r3 = bm_radius(u3);
    const float uz = r1 * cos_turn(u2), uy = r1 * sin_turn(u2), ux = r3 * cos_turn(u4);
    const float4 c = v[2], ca = v[4], cb = v[5];
    const float wx = c.x + ca.x * ux;
    const float wy = c.y + (ca.y * ux + ca.z * uy);
    const float wz = c.z + (cb.x * ux + cb.y * uy + cb.z * uz);
    const float px = wx - f.pos[0], py = wy - f.pos[1], pz = wz - f.pos[2];
    const float camz = dot3(f.rot[2], px, py, pz);
    const float dz = camz - f.near_z, mz = 1.001f * b.z;
    if (!(dz > mz)) return dz <= -mz ? 0 : -1;
    if (!(f.flags & FLAG_RADIAL)) return 1;
    uint32_t c5 = t0, c6 = t1, ck = t2;   // the rest of the chain
    if (!(v[3].w < 0.0f)) {
        uint32_t c = t0;
        if (f.flags & FLAG_JITTER) {
            c5 = pcg(c);
            c6 = pcg(c5);
            c = c6;
        }
        ck = pcg(c);
    }
    const float zl = camz - b.z;   // <= camz on both paths
    if (!(zl > 0.0f)) return -1;
    const float camx = dot3(f.rot[0], px, py, pz), camy = dot3(f.rot[1], px, py, pz);
    const float iz = rcp(camz), izl = rcp(zl);
    const float t1x = f.intr[0] * camx, t1y = f.intr[1] * camy;
    const float t2x = t1x * iz, t2y = t1y * iz;
    float pixx = t2x + f.intr[2];
    float pixy = t2y + f.intr[3];
    float jb = 0.0f;
    if (f.flags & FLAG_JITTER) {
        const float jy = r3 * sin_turn(u4);
        const float u5 = fmax(unif24(c5), 5.96e-8f), u6 = unif24(c6);
        const float jx = bm_radius(u5) * cos_turn(u6);
        pixx = pixx + f.jitter * jx;
        pixy = pixy + f.jitter * jy;
        jb = f.jitter * (DELTA_N + 6.0f * EPS * U_MAX);
    }
    const float4 a = v[0], d = v[3];
    // |dv - emit_pbvr's dv|: the projection bound, then 8 EPS |t2| + E_RCP |t2| + 6 EPS |c| + 2 EPS |a| for the
    // roundings of the pixel, the jittered sum and dv on both paths and v_rcp_f32, and the jitter's part
    const float bx = ((fabsf(f.intr[0]) * b.x + 2.0f * EPS * fabsf(t1x) + fabsf(t2x) * b.z) * izl +
                      (8.0f * EPS + E_RCP) * fabsf(t2x) + 6.0f * EPS * fabsf(f.intr[2]) +
                      2.0f * EPS * fabsf(a.x) + jb) * 1.01f;
    const float by = ((fabsf(f.intr[1]) * b.y + 2.0f * EPS * fabsf(t1y) + fabsf(t2y) * b.z) * izl +
                      (8.0f * EPS + E_RCP) * fabsf(t2y) + 6.0f * EPS * fabsf(f.intr[3]) +
                      2.0f * EPS * fabsf(a.y) + jb) * 1.01f;
    const float dvx = pixx - a.x, dvy = pixy - a.y;
    const float r2 = c.w * dvx * dvx + 2.0f * d.x * dvx * dvy + d.y * dvy * dvy;
    const float ex = fabsf(dvx), ey = fabsf(dvy), q0 = fabsf(c.w), q1 = fabsf(d.x), q2 = fabsf(d.y);
    const float gx = q0 * ex + q1 * ey, gy = q1 * ex + q2 * ey;
    const float cq = q0 * bx * bx + 2.0f * q1 * bx * by + q2 * by * by;
    const float br2 = (2.0f * (gx * bx + gy * by) + cq + 10.0f * EPS * (gx * ex + gy * ey)) * 1.01f;
    if (!(br2 <= 200.0f)) return -1;   // (and NaN)
    const float oo = fmin(a.w, 1.0f - 1e-6f);
    const float t = exp_half_neg(fmax(r2, 0.0f));
    const float y = oo * t;
    const float tf = (t + TAU_F) * (1.0f + 2.0f * ETA_TF);   // >= e^(-r2/2) at this path's r2
    // e^(br2/2) <= 1 + 0.65 br2 for br2 <= 1, and <= 2^ceil(0.7214 br2) (0.7214 >= log2(e) / 2) above
    const float grow = br2 <= 1.0f ? 1.0f + 0.65f * br2 : __builtin_amdgcn_ldexpf(1.0f, (int)ceilf(0.7214f * br2));
    const float te = tf * grow;                               // >= e^(-r2/2) at emit_pbvr's r2
    const float dt = te * (0.5f * br2 + ETA_TE) + ETA_TF * tf + (TAU_E + TAU_F);
    const float dy = (oo * dt + 2.0f * EPS * y) * 1.01f;     // >= |y - emit_pbvr's y|
    const float x = fmin(y, 1.0f - 1e-6f);
    const float xmax = fmin(y + dy, 1.0f - 1e-6f), xmin = fmax(x - dy, 0.0f);
    const float h = extinction_over_linear(x), rdz = rcp(d.z);
    const float kf = clamp(h * rdz, 0.0f, 1.0f);
    const float dh = dy * rcp(1.0f - xmax) * 1.001f;         // >= |h(x) - h(emit_pbvr's x)|
    const float rx = rcp(xmin);                               // +inf for xmin = 0
    const float eta = fmin(ETA_HE, ETA_HE_X * rx) + fmin(ETA_HF, ETA_HF_X * rx);
    const float hu = h * (1.0f + 2.0f * ETA_HF) + dh;        // >= h at both points
    const float dk = (dh + (eta + 2.0f * EPS + E_RCP) * hu) * rdz * 1.01f + 4.0f * EPS;
    const float uk = unif24(ck);
    if (uk < kf - dk) return 1;
    if (uk >= kf + dk) return 0;
    return -1;
}

// OUT = 0: packed 64-bit minimum into the depth|id buffer of the particle's ensemble.
// OUT = 1: (single ensemble) per-particle key (depth key << 32 | particle index) and subpixel into pkeys /
//          pidx at the particle's index, for the depth-sorted resolve.
// Occupancy target of the PBVR packed-minimum path: seven waves per SIMD at 256 threads (the LDS below, 22.3 KB at
// 4 particles per thread, lets a CU hold seven blocks; the register limit costs a few bytes of spills, which the
// count-only views repay), six at 128 and 512 threads, where the exact path's spills would cost more (measured).
// The other paths keep the compiler's choice.
// The body run by block `rank` of the `nblocks` blocks that share one view's nens ensembles (k_particles: the whole
// grid; k_particles_fused: the blocks a batch round assigns to one item).
template <int MK, int PPT, int OUT, int BS>
GPS_DEV void particles_body(const GpsFrame &f, const float4 *__restrict__ view, const uint32_t *__restrict__ offsets,
                            const uint32_t *__restrict__ totals, uint32_t nens, uint32_t first_seed,
                            uint64_t *__restrict__ depth, uint64_t sub_count, const float *__restrict__ lut,
                            uint64_t *__restrict__ stats, uint64_t *__restrict__ pkeys, uint32_t *__restrict__ pidx,
                            const uint32_t *__restrict__ perm, uint32_t rank, uint32_t nblocks) {
    constexpr uint32_t CH = BS * PPT;
    // PBVR packed-minimum path: a particle of a splat marked off-screen has its PCG chain drawn here; it is
    // settled when the draws alone decide (the view-conditioned draw rejects, or, for a splat whose particles all
    // pass the near plane, there is no radial correction or the keep draw is below the splat's lower bound of the
    // keep ratio, -v[3].w).  Otherwise its record (splat and chain) goes to its wave's work buffer, which runs
    // count_only on whole waves of 64 records as soon as it holds 64; records count_only cannot decide go to
    // the wave's emit_pbvr queue, run on whole waves of 64 the same way.  The remainders run at the end.
    constexpr bool COUNT_ONLY = MK == 1 && OUT == 0;
    constexpr uint32_t NWAVE = BS / 64;
    constexpr uint32_t WCAP = COUNT_ONLY ? PARTICLE_WAVE_CAP : 1u;   // per wave: fewer than 64 waiting + 64 appended
    constexpr int NV = MK == 0 ? 2 : (MK == 2 ? 4 : VIEW_VEC4);   // view record float4s the emitter reads
    __shared__ uint32_t s_cstart[MAX_BATCH + 1];
    constexpr uint32_t OCAP = 256u;   // chunk splat ranges up to this many positions are staged in LDS
    constexpr uint32_t CRUN = 8u;     // consecutive chunks per block turn
    __shared__ uint32_t s_lo, s_hi;
    __shared__ uint32_t s_qpeak, s_wpeak;   // high-water marks of a wave's queue and work buffer (ST_QPEAK, ST_WPEAK)
    __shared__ uint32_t s_off[OCAP + 1], s_prm[OCAP];
    __shared__ uint64_t s_pts[NWAVE], s_hits[NWAVE];
    // work records: (splat, c1, c2, c3) and (c4, t0, t1, t2) of count_only; queue entries: (splat, seed)
    __shared__ uint4 s_wa[COUNT_ONLY ? NWAVE : 1][WCAP], s_wb[COUNT_ONLY ? NWAVE : 1][WCAP];
    __shared__ uint2 s_eq[COUNT_ONLY ? NWAVE : 1][WCAP];
    static_assert(PPT == 1 || PPT == 2 || PPT == 4 || PPT == 8, "particles per thread");
    static_assert(BS == 128 || BS == 256 || BS == 512, "workgroup size");
    const bool count_only_on = COUNT_ONLY && (f.flags & FLAG_FAST) != 0u;
    if (threadIdx.x == 0) {
        s_qpeak = 0u;
        s_wpeak = 0u;
        uint32_t acc = 0;
        for (uint32_t b = 0; b < nens; b++) {
            s_cstart[b] = acc;
            acc += (uint32_t)(((uint64_t)totals[b] + (CH - 1u)) / CH);   // 64-bit: totals reach 2^32 - 1
        }
        s_cstart[nens] = acc;
    }
    __syncthreads();
    const uint32_t nchunks = s_cstart[nens];
    const uint32_t N = f.N;
    // generated points (PBVR / occupancy statistic) in the low 32 bits, particles that reach the packed minimum or
    // write a subpixel in the high 32 bits; a thread's share of a launch stays below 2^32 particles (at most
    // 64 ensembles of fewer than 2^32 particles over at least 32,768 threads)
    uint64_t pts = 0;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // this wave's pending count_only records [0, rn) and emit_pbvr entries [0, qn) (the same in every lane)
    uint32_t rn = 0, qn = 0;
    // LDS written by other lanes of the wave is read after this
    auto wave_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // the first n entries of the wave's queue through emit_pbvr, one per lane; the rest (n = 64) to the front.
    // All are of splats marked off-screen: emit_pbvr decides their count, and none of their particles can reach
    // a subpixel (the box test that marked the splat), so nothing is written.
    auto run_queue = [&](uint32_t n) {
        wave_sync();
        if (lane < n) {
            const uint2 e = s_eq[wave][lane];
            Rng r(e.y);
            pts += emit_pbvr(f, r, view + (size_t)VIEW_VEC4 * e.x).counted;
        }
        qn -= n;
        if (lane < qn) s_eq[wave][lane] = s_eq[wave][n + lane];   // n = 64 > qn: no overlap
        wave_sync();
    };
    // the first n records of the work buffer through count_only, one per lane; the rest (n = 64) to the front
    auto run_work = [&](uint32_t n, float4 *v, uint32_t &vsp) {
        wave_sync();
        bool undecided = false;
        uint32_t qs = 0, qseed = 0;
        if (lane < n) {
            const uint4 a = s_wa[wave][lane], b = s_wb[wave][lane];
            if (a.x != vsp) {
#pragma unroll
                for (int q = 0; q < NV; q++) v[q] = view[(size_t)VIEW_VEC4 * a.x + q];
                vsp = a.x;
            }
            const int c = count_only(f, v, a.y, a.z, a.w, b.x, b.y, b.z, b.w);
            if (c >= 0) {
                pts += (uint32_t)c;
            } else {
                undecided = true;
                qs = a.x;
                qseed = pcg_inv(a.y);   // the particle's seed: c1 = pcg(seed)
            }
        }
        rn -= n;
        if (lane < rn) {   // n = 64 > rn: no overlap
            s_wa[wave][lane] = s_wa[wave][n + lane];
            s_wb[wave][lane] = s_wb[wave][n + lane];
        }
        const uint64_t m = __ballot(undecided);
        if (undecided) {
            const uint32_t k = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            s_eq[wave][qn + k] = make_uint2(qs, qseed);
        }
        qn += (uint32_t)__popcll(m);
        if (lane == 0u && qn > 0u) atomicMax(&s_qpeak, qn);
        wave_sync();
        if (qn >= 64u) run_queue(64u);
    };
    uint32_t prev = 0xFFFFFFFFu, prev_hi = 0;   // the block's previous chunk and the owner of its last particle
    // runs of CRUN consecutive chunks per block: a chunk's owners are searched from the previous chunk's last
    for (uint32_t chunk = rank * CRUN; chunk < nchunks;
         chunk = (chunk % CRUN == CRUN - 1u) ? chunk + (nblocks - 1u) * CRUN + 1u : chunk + 1u) {
        uint32_t b = 0;
        while (b + 1u < nens && s_cstart[b + 1u] <= chunk) b++;
        const uint32_t total = totals[b];
        const uint32_t p0 = (chunk - s_cstart[b]) * CH;
        const uint32_t pend = (uint32_t)min((uint64_t)p0 + CH, (uint64_t)total);   // p0 + CH may pass 2^32
        const uint32_t *offs = offsets + (size_t)b * N;
        uint64_t *dep = depth + (size_t)b * sub_count;
        const uint32_t seed_mix = (first_seed + b) * 2654435761u + 1u;
        const bool follows = prev + 1u == chunk && chunk > s_cstart[b];   // the same ensemble's next chunk
        if (threadIdx.x < 128u) {   // waves 0 and 1: the owners of the chunk's first and last particles
            const uint32_t s = owner_search_wave(offs, follows ? prev_hi : 0u, N - 1u,
                                                 threadIdx.x < 64u ? p0 : pend - 1u, follows);
            if (threadIdx.x == 0) s_lo = s;
            else if (threadIdx.x == 64u) s_hi = s;
        }
        __syncthreads();
        prev = chunk;
        prev_hi = s_hi;
        // the chunk's positions s_lo..s_hi: their first particles (and the next position's, or the total) and
        // splats, staged in LDS when a search among them would take more than a few steps and they fit
        const uint32_t lo = s_lo, R = s_hi - lo + 1u;
        const bool staged = R > 16u && R <= OCAP;
        if (staged) {
            for (uint32_t t = threadIdx.x; t <= R; t += BS) {
                const uint32_t gi = lo + t;
                s_off[t] = gi < N ? offs[gi] : total;
                if (t < R) s_prm[t] = perm ? perm[gi] : gi;
            }
            __syncthreads();
        }
        // this thread's particles are p0 + q for q in [q0, min(q0 + PPT, cnt)); the comparisons use q, which
        // cannot wrap (p0 + q0 can, in the last chunk of an ensemble with more than 2^32 - CH particles)
        const uint32_t cnt = pend - p0, q0 = threadIdx.x * PPT;
        const uint32_t pfirst = p0 + q0;
        // position in the splat order of the offsets, its first particle, the next position's and the splat
        uint32_t i = 0, cur = 0, nxt = 0, sp = 0;
        if (q0 < cnt) {
            if (staged) {
                uint32_t a = 0, z = R - 1u;
                while (a < z) {
                    const uint32_t mid = (a + z + 1u) >> 1;
                    if (s_off[mid] <= pfirst) a = mid; else z = mid - 1u;
                }
                i = lo + a;
                cur = s_off[a];
                nxt = s_off[a + 1u];
                sp = s_prm[a];
            } else {
                i = owner_search(offs, lo, s_hi, pfirst);
                cur = offs[i];
                nxt = (i + 1u < N) ? offs[i + 1u] : total;
                sp = perm ? perm[i] : i;
            }
        }
        // one particle's emit and output, the splat's view record in v
        auto emit_out = [&](Rng &r, uint32_t p, uint32_t k, uint32_t s2, const float4 *v) {
            Hit h;
            if constexpr (MK == 0) h = emit_gps(f, r, v[0], v[1], lut);
            else if constexpr (MK == 1) h = emit_pbvr(f, r, v);
            else h = emit_occupancy(f, r, k, v);
            pts += (uint64_t)h.counted + ((uint64_t)(h.idx != NO_HIT) << 32);
            if (OUT == 0) {
                if (h.idx != NO_HIT) write_min(dep + h.idx, ((uint64_t)h.key << 32) | (uint64_t)s2);
            } else {
                pkeys[p] = ((uint64_t)h.key << 32) | (uint64_t)p;
                pidx[p] = h.idx;
            }
        };
        if constexpr (!COUNT_ONLY) {
            if (q0 < cnt) {
                // the splat's view record stays in registers while its particles run (loaded when the splat
                // changes, so its latency overlaps the seed hashing instead of stalling each particle)
                float4 v[NV];
#pragma unroll
                for (int q = 0; q < NV; q++) v[q] = view[(size_t)VIEW_VEC4 * sp + q];
#pragma unroll 1
                for (uint32_t j = 0; j < PPT; j++) {
                    if (q0 + j >= cnt) break;
                    const uint32_t p = pfirst + j;
                    if (p >= nxt) {   // every particle of the chunk belongs to a position <= s_hi
                        do {
                            i++;
                            cur = nxt;
                            nxt = staged ? s_off[i + 1u - lo] : ((i + 1u < N) ? offs[i + 1u] : total);
                        } while (p >= nxt);
                        sp = staged ? s_prm[i - lo] : (perm ? perm[i] : i);
#pragma unroll
                        for (int q = 0; q < NV; q++) v[q] = view[(size_t)VIEW_VEC4 * sp + q];
                    }
                    const uint32_t k = p - cur;
                    Rng r(particle_seed_mixed(sp, k + 1u, seed_mix));
                    emit_out(r, p, k, sp, v);
                }
            }
        } else {
            // the view record of splat vsp stays in registers while its particles run
            float4 v[NV];
            uint32_t vsp = 0xFFFFFFFFu;
            auto load_view = [&](uint32_t s2) {
                if (s2 != vsp) {
#pragma unroll
                    for (int q = 0; q < NV; q++) v[q] = view[(size_t)VIEW_VEC4 * s2 + q];
                    vsp = s2;
                }
            };
            if (q0 < cnt) load_view(sp);
#pragma unroll 1
            for (uint32_t j = 0; j < PPT; j++) {   // every lane runs PPT iterations (the work buffer needs whole waves)
                const uint32_t p = pfirst + j;
                bool push = false;
                uint32_t c1, c2, c3, c4, t0, t1, t2;   // the record; set on every path that pushes it
                if (q0 + j < cnt) {
                    if (p >= nxt) {   // every particle of the chunk belongs to a position <= s_hi
                        do {
                            i++;
                            cur = nxt;
                            nxt = staged ? s_off[i + 1u - lo] : ((i + 1u < N) ? offs[i + 1u] : total);
                        } while (p >= nxt);
                        sp = staged ? s_prm[i - lo] : (perm ? perm[i] : i);
                    }
                    load_view(sp);
                    const uint32_t k = p - cur;
                    const uint32_t seed = particle_seed_mixed(sp, k + 1u, seed_mix);
                    if (count_only_on && v[3].w != 0.0f) {
                        // the chain of emit_pbvr's draws: c1..c4 the normals, the view-conditioned draw, then (for
                        // a splat whose particles all pass the near plane, with the radial correction) the jitter
                        // pair and the keep draw
                        c1 = pcg(seed);
                        c2 = pcg(c1);
                        c3 = pcg(c2);
                        c4 = pcg(c3);
                        uint32_t c = c4;
                        int dec = -1;
                        if (f.method == M_VIEW_CONDITIONED) {
                            c = pcg(c);
                            if (unif24(c) >= v[1].w) dec = 0;
                        }
                        if (!(v[3].w < 0.0f)) {   // count_only hashes the rest when it needs it
                            t0 = t1 = t2 = c;
                        } else if (dec < 0) {
                            if (!(f.flags & FLAG_RADIAL)) {
                                dec = 1;
                            } else {
                                t0 = t1 = c;
                                if (f.flags & FLAG_JITTER) {
                                    t0 = pcg(c);
                                    t1 = pcg(t0);
                                    c = t1;
                                }
                                t2 = pcg(c);
                                if (unif24(t2) < -v[3].w) dec = 1;   // below every particle's keep ratio
                            }
                        }
                        if (dec >= 0) pts += (uint32_t)dec;
                        else push = true;
                    } else {
                        Rng r(seed);
                        emit_out(r, p, k, sp, v);
                    }
                }
                if (count_only_on) {   // every lane of the wave is here
                    const uint64_t m = __ballot(push);
                    if (push) {
                        const uint32_t at = rn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        s_wa[wave][at] = make_uint4(sp, c1, c2, c3);
                        s_wb[wave][at] = make_uint4(c4, t0, t1, t2);
                    }
                    rn += (uint32_t)__popcll(m);   // at most 63 + 64 < WCAP
                    if (lane == 0u && rn > 0u) atomicMax(&s_wpeak, rn);
                    if (rn >= 64u) run_work(64u, v, vsp);
                }
            }
        }
        __syncthreads();   // the chunk's owners and staged offsets are read by every wave before the next chunk
    }
    if (count_only_on) {   // the remainders (fewer than 64 each), one per lane
        float4 v[NV];
        uint32_t vsp = 0xFFFFFFFFu;
        if (rn) run_work(rn, v, vsp);
        if (qn) run_queue(qn);
    }
    {
        const uint64_t sp = wave_sum(pts & 0xFFFFFFFFull), sh = wave_sum(pts >> 32);
        if ((threadIdx.x & 63) == 0) {
            s_pts[threadIdx.x >> 6] = sp;
            s_hits[threadIdx.x >> 6] = sh;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = 0, th = 0;
            for (int w = 0; w < BS / 64; w++) {
                t += s_pts[w];
                th += s_hits[w];
            }
            if (MK != 0 && t) atomicAdd((unsigned long long *)&stats[ST_POINTS], (unsigned long long)t);
            if (th) atomicAdd((unsigned long long *)&stats[ST_HITS], (unsigned long long)th);
            if (COUNT_ONLY) {   // the buffers' high-water marks (an atomic only when above the recorded one)
                if (s_qpeak > __hip_atomic_load(&stats[ST_QPEAK], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMax((unsigned long long *)&stats[ST_QPEAK], (unsigned long long)s_qpeak);
                if (s_wpeak > __hip_atomic_load(&stats[ST_WPEAK], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicMax((unsigned long long *)&stats[ST_WPEAK], (unsigned long long)s_wpeak);
            }
        }
    }
}

template <int MK, int PPT, int OUT, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(MK == 1 && OUT == 0 ? (BS == 256 ? 7 : 6) : 1)))
void k_particles(GpsFrame f, const float4 *__restrict__ view, const uint32_t *__restrict__ offsets,
                 const uint32_t *__restrict__ totals, uint32_t nens, uint32_t first_seed,
                 uint64_t *__restrict__ depth, uint64_t sub_count, const float *__restrict__ lut,
                 uint64_t *__restrict__ stats, uint64_t *__restrict__ pkeys, uint32_t *__restrict__ pidx,
                 const uint32_t *__restrict__ perm) {
    particles_body<MK, PPT, OUT, BS>(f, view, offsets, totals, nens, first_seed, depth, sub_count, lut, stats, pkeys,
                                     pidx, perm, blockIdx.x, gridDim.x);
}

// ================================================================== splat-parallel GPS path
// gps_splat.wgsl:304-333: one thread per splat, one sequential stream for its count and particles
__global__ __launch_bounds__(64) void k_splat_pass(GpsFrame f, const float4 *__restrict__ view,
                                                   const uint2 *__restrict__ cinfo, uint32_t seed,
                                                   uint64_t *__restrict__ depth, const float *__restrict__ lut,
                                                   uint64_t *__restrict__ stats) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f.N) return;
    uint2 ci = cinfo[i];
    if ((ci.y & 3u) == KIND_NONE) return;
    Rng r(particle_seed(i, 0u, seed));
    uint32_t n = poisson(r, gpsm::fbits(ci.x));
    atomicAdd((unsigned long long *)&stats[ST_POINTS], (unsigned long long)n);
    atomicAdd((unsigned long long *)&stats[ST_VISIBLE], 1ull);
    if (n > f.cap) {
        atomicAdd((unsigned long long *)&stats[ST_TRUNC], 1ull);
        atomicAdd((unsigned long long *)&stats[ST_DROPPED], (unsigned long long)(n - f.cap));
        n = f.cap;
    }
    const float4 *v = view + (size_t)VIEW_VEC4 * i;
    float4 v0 = v[0], v1 = v[1];
    for (uint32_t k = 0; k < n; k++) {
        Hit h = emit_gps(f, r, v0, v1, lut);
        if (h.idx != NO_HIT) write_min(depth + h.idx, ((uint64_t)h.key << 32) | (uint64_t)i);
    }
}

// ================================================================== resolve
// gps_splat.wgsl:361-386, for every ensemble of the batch in ensemble order
// With `mom` set it also accumulates, per pixel, the sum of squares of each ensemble's pixel value (x, y, z)
// and the number of covered subpixels (w), and keeps in `zmin` the nearest depth seen over the ensembles:
// the denoiser's variance, particle-count and depth features.
// The body for pixel block `blk` of one view (k_resolve; k_resolve_fused: every item of a batch round).
GPS_DEV void resolve_body(const GpsFrame &f, uint64_t *__restrict__ depth, uint64_t sub_count,
                          const uint32_t *__restrict__ skip, uint32_t nens, const float4 *__restrict__ colour,
                          float4 *__restrict__ accum, int clear, uint64_t *__restrict__ stats, float4 *__restrict__ mom,
                          float *__restrict__ zmin, uint32_t blk) {
    const uint32_t X = f.W * f.H;
    const uint32_t spp = f.spp_side * f.spp_side;
    uint32_t i = blk * blockDim.x + threadIdx.x;
    if (i == 0) atomicAdd((unsigned long long *)&stats[ST_ENSEMBLES], (unsigned long long)nens);
    if (i >= X) return;
    float4 acc = accum[i];
    float4 m2 = mom ? mom[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t zbits = mom ? __float_as_uint(zmin[i]) : 0u;
    const float fs = (float)spp;
    for (uint32_t b = 0; b < nens; b++) {
        uint64_t *d = depth + (size_t)b * sub_count + (size_t)i * spp;
        if (skip == nullptr || skip[b] == 0u) {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            uint32_t hits = 0u;
            for (uint32_t k = 0; k < spp; k++) {
                uint64_t w = d[k];
                if (w == EMPTY64) {
                    sx = sx + f.bg[0];
                    sy = sy + f.bg[1];
                    sz = sz + f.bg[2];
                } else {
                    float4 c = colour[(uint32_t)w];
                    sx = sx + c.x;
                    sy = sy + c.y;
                    sz = sz + c.z;
                    hits++;
                    // the key is the depth's bits with the sign bit set: unsigned order is depth order
                    const uint32_t zb = (uint32_t)(w >> 32) & 0x7FFFFFFFu;
                    zbits = zb < zbits ? zb : zbits;
                }
            }
            const float vx = sx / fs, vy = sy / fs, vz = sz / fs;
            acc.x = acc.x + vx;
            acc.y = acc.y + vy;
            acc.z = acc.z + vz;
            acc.w = acc.w + 1.0f;
            if (mom) {
                m2.x = m2.x + vx * vx;
                m2.y = m2.y + vy * vy;
                m2.z = m2.z + vz * vz;
                m2.w = m2.w + (float)hits;
            }
        }
        if (clear)
            for (uint32_t k = 0; k < spp; k++) d[k] = EMPTY64;
    }
    accum[i] = acc;
    if (mom) {
        mom[i] = m2;
        zmin[i] = __uint_as_float(zbits);
    }
}

__global__ __launch_bounds__(256) void k_resolve(GpsFrame f, uint64_t *__restrict__ depth, uint64_t sub_count,
                                                 const uint32_t *__restrict__ skip, uint32_t nens,
                                                 const float4 *__restrict__ colour, float4 *__restrict__ accum,
                                                 int clear, uint64_t *__restrict__ stats, float4 *__restrict__ mom,
                                                 float *__restrict__ zmin) {
    resolve_body(f, depth, sub_count, skip, nens, colour, accum, clear, stats, mom, zmin, blockIdx.x);
}

// moments to zero and nearest depth to +inf
__global__ void k_features_clear(float4 *__restrict__ mom, float *__restrict__ zmin, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        mom[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        zmin[i] = __uint_as_float(0x7F800000u);
    }
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, uint64_t n, uint64_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_stats_add(uint64_t *stats, int slot, uint64_t v) {
    if (threadIdx.x == 0) stats[slot] += v;
}

__global__ void k_stats_clear(uint64_t *stats, int first, int count) {
    if ((int)threadIdx.x < count) stats[first + threadIdx.x] = 0ull;
}

__global__ void k_copy_f4(float4 *__restrict__ dst, const float4 *__restrict__ src, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_accum_clear(float4 *__restrict__ a, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ================================================================== composite
// gps_composite.wgsl:153-179 into packed RGBA8 (R in the low byte)
GPS_DEV float oetf(float c) {
    if (c <= 0.0031308f) return 12.92f * c;
    return 1.055f * gpsm::exp((1.0f / 2.4f) * gpsm::log(c)) - 0.055f;
}

__global__ __launch_bounds__(256) void k_composite(const float4 *__restrict__ accum, uint32_t W, uint32_t H,
                                                   float exposure, uint32_t gamma_on, uint32_t ow, uint32_t oh,
                                                   uint32_t *__restrict__ out) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ow * oh) return;
    uint32_t X = t % ow, Y = t / ow;
    float uvx = ((float)X + 0.5f) / (float)ow, uvy = ((float)Y + 0.5f) / (float)oh;
    float qx = uvx * (float)W - 0.5f, qy = uvy * (float)H - 0.5f;
    float flx = floorf(qx), fly = floorf(qy);
    int px = (int)flx, py = (int)fly;
    float fx = qx - flx, fy = qy - fly;
    float lin[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int xx = px + (k & 1), yy = py + (k >> 1);
        xx = xx < 0 ? 0 : (xx > (int)W - 1 ? (int)W - 1 : xx);
        yy = yy < 0 ? 0 : (yy > (int)H - 1 ? (int)H - 1 : yy);
        float4 a = accum[(size_t)yy * W + (size_t)xx];
        float wgt = fmax(a.w, 1.0f);
        lin[k][0] = a.x / wgt;
        lin[k][1] = a.y / wgt;
        lin[k][2] = a.z / wgt;
    }
    uint32_t packed = 0xff000000u;
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float top = mix(lin[0][ch], lin[1][ch], fx);
        float bot = mix(lin[2][ch], lin[3][ch], fx);
        float c = mix(top, bot, fy) * exposure;
        c = clamp(c, 0.0f, 1.0f);
        if (gamma_on) c = oetf(c);
        packed |= (f2u(gpsm::rint(c * 255.0f)) & 0xffu) << (8 * ch);
    }
    out[t] = packed;
}

// ================================================================== phase-0 unit kernels
// shaders/p0_sample.wgsl:205-243 with the two passes collapsed into one packed 64-bit minimum:
// (depth key << 32) | colour.  The colour tie rule (smallest colour among equal keys) is the
// low word's order.
__global__ __launch_bounds__(256) void k_phase0(uint32_t seed, uint32_t pixels, uint32_t samples,
                                                uint64_t *__restrict__ buf) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= samples) return;
    uint32_t h0 = pcg(i ^ pcg(seed)), h1 = pcg(h0), h2 = pcg(h1);
    uint32_t pix = h0 % pixels;
    uint32_t key = ((0x3F800000u | (h1 & 0x007FFFFFu)) | 0x80000000u) & 0xFFFF0000u;
    uint32_t rgb = h2 & 0x00FFFFFFu;
    __hip_atomic_fetch_min(&buf[pix], ((uint64_t)key << 32) | rgb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// exclusive u32 scan of a plain array with the same tile / look-back machinery as k_count_scan
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_u32(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                         uint64_t n, uint32_t *ticket, uint64_t *status,
                                                         uint32_t tiles, uint32_t epoch) {
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_wave[SCAN_BLOCK / 64];
    __shared__ uint64_t s_prefix;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint64_t tsum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = base + j < n ? in[base + j] : 0u;
        tsum += v[j];
    }
    uint64_t texcl;
    uint64_t agg = block_exclusive_scan(tsum, texcl, s_wave);
    if (threadIdx.x == 0) s_prefix = lookback<true>(status, tile, agg, epoch);
    __syncthreads();
    uint64_t run = s_prefix + texcl;   // the output keeps the low 32 bits
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        if (base + j < n) out[base + j] = (uint32_t)run;
        run += v[j];
    }
    if (tile == tiles - 1u && threadIdx.x == 0) *ticket = 0u;
}

__global__ void k_math(int fn, const float *__restrict__ x, const float *__restrict__ y, float *__restrict__ out,
                       uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = x[i], r;
    switch (fn) {
    case 0: r = gpsm::exp(a); break;
    case 1: r = gpsm::log(a); break;
    case 2: r = gpsm::sin(a); break;
    case 3: r = gpsm::cos(a); break;
    case 4: r = gpsm::sqrt(a); break;
    default: r = a / y[i]; break;
    }
    out[i] = r;
}


// ================================================================== fused batch (gps_batch_render)
// Every launch covers all items of a batch; each block finds its item (and its block index within the item) in a
// table the host builds at record time, or, for the particle kernel, one k_fused_assign writes on the device once
// the round's particle totals are known.  The bodies are the single-view kernels' own.  A round's ensembles of
// item i are [ens0, ens0 + nens[i]) of the item's own sequence: nens comes from the host's table (0 once the
// item's ensembles are done), ens0 is the round's first ensemble (added to the item's first seed modulo 2^32, as
// the single-view path adds the ensembles it has done).

// grid = (blocks over the largest image, items): each item's image, denoiser moments (nearest depth to +inf),
// counters, and the count-only splat order's counter word, before the batch
__global__ __launch_bounds__(256) void k_fused_clear(const FusedItem *__restrict__ items) {
    const FusedItem &it = items[blockIdx.y];
    const uint64_t X = (uint64_t)it.f.W * it.f.H;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < X; i += (uint64_t)gridDim.x * blockDim.x) {
        it.accum[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (it.mom) {
            it.mom[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            it.zmin[i] = __uint_as_float(0x7F800000u);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (uint32_t)ST_WORDS) it.stats[threadIdx.x] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < 2u && it.perm) it.perm[perm_counter(it.f.N) + threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void k_view_fused(const FusedItem *__restrict__ items,
                                                    const uint32_t *__restrict__ blk_item,
                                                    const uint32_t *__restrict__ item_blk0) {
    const uint32_t item = blk_item[blockIdx.x];
    const FusedItem &it = items[item];
    view_body(it.f, it.splats, it.sh, it.view, it.cinfo, it.colour, it.perm, blockIdx.x - item_blk0[item]);
}

// blk_pair: item << 8 | ensemble of the round, one entry per (item, ensemble) tile
__global__ __launch_bounds__(SCAN_BLOCK) void k_count_scan_fused(const FusedItem *__restrict__ items,
                                                                const uint32_t *__restrict__ blk_pair, uint32_t ens0,
                                                                uint32_t epoch) {
    const uint32_t pr = blk_pair[blockIdx.x];
    const FusedItem &it = items[pr >> 8];
    count_scan_body(it.f, it.cinfo, it.first_seed + ens0, it.offsets, it.bs, it.tiles, epoch, it.stats, it.perm,
                    pr & 255u);
}

// One block: gives each item of emitter class mk that has particles this round a contiguous range of the particle
// grid, at least one block and otherwise in proportion to its chunks of `chunk` particles; blocks beyond the
// ranges get FUSED_NONE.  nitems <= FUSED_MAX_ITEMS = blockDim.x, grid >= nitems.
__global__ __launch_bounds__(FUSED_MAX_ITEMS) void k_fused_assign(const FusedItem *__restrict__ items,
                                                                  uint32_t nitems, const uint32_t *__restrict__ nens,
                                                                  uint32_t mk, uint32_t chunk, uint32_t grid,
                                                                  uint32_t *__restrict__ blk_item,
                                                                  uint32_t *__restrict__ blk_rank,
                                                                  uint32_t *__restrict__ item_nblk) {
    __shared__ uint32_t s_off[FUSED_MAX_ITEMS + 1], s_nb[FUSED_MAX_ITEMS];
    __shared__ unsigned long long s_total;
    __shared__ uint32_t s_active;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        s_total = 0ull;
        s_active = 0u;
    }
    __syncthreads();
    uint64_t ch = 0;
    if (t < nitems && items[t].mk == mk) {
        const FusedItem &it = items[t];
        const uint32_t ne = nens[t];
        for (uint32_t e = 0; e < ne; e++) ch += ((uint64_t)it.bs.totals[e] + (chunk - 1u)) / chunk;
    }
    if (ch) {
        atomicAdd(&s_total, (unsigned long long)ch);
        atomicAdd(&s_active, 1u);
    }
    __syncthreads();
    const uint64_t T = s_total;
    const uint32_t A = s_active;
    const uint32_t nb = ch ? 1u + (uint32_t)((uint64_t)(grid - A) * ch / T) : 0u;   // sum over items <= grid
    if (t < nitems) {
        s_nb[t] = nb;
        item_nblk[t] = nb;
    }
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (uint32_t i = 0; i < nitems; i++) {
            s_off[i] = acc;
            acc += s_nb[i];
        }
        s_off[nitems] = acc;
    }
    __syncthreads();
    const uint32_t used = s_off[nitems];
    for (uint32_t g = t; g < grid; g += blockDim.x) {
        uint32_t item = FUSED_NONE, rank = 0;
        if (g < used) {   // the last item whose range starts at or before g (items without blocks have empty ranges)
            uint32_t lo = 0, hi = nitems - 1u;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (s_off[mid] <= g) lo = mid; else hi = mid - 1u;
            }
            item = lo;
            rank = g - s_off[lo];
        }
        blk_item[g] = item;
        blk_rank[g] = rank;
    }
}

template <int MK, int PPT, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(MK == 1 ? (BS == 256 ? 7 : 6) : 1)))
void k_particles_fused(const FusedItem *__restrict__ items, const uint32_t *__restrict__ nens, uint32_t ens0,
                       const uint32_t *__restrict__ blk_item, const uint32_t *__restrict__ blk_rank,
                       const uint32_t *__restrict__ item_nblk) {
    const uint32_t item = blk_item[blockIdx.x];
    if (item == FUSED_NONE) return;   // the whole block
    const FusedItem &it = items[item];
    particles_body<MK, PPT, 0, BS>(it.f, it.view, it.offsets, it.bs.totals, nens[item], it.first_seed + ens0,
                                   it.depth, it.sub_count, it.lut, it.stats, nullptr, nullptr, it.perm,
                                   blk_rank[blockIdx.x], item_nblk[item]);
}

__global__ __launch_bounds__(256) void k_resolve_fused(const FusedItem *__restrict__ items,
                                                       const uint32_t *__restrict__ nens,
                                                       const uint32_t *__restrict__ blk_item,
                                                       const uint32_t *__restrict__ item_blk0) {
    const uint32_t item = blk_item[blockIdx.x];
    const FusedItem &it = items[item];
    const uint32_t ne = nens[item];
    if (ne == 0) return;
    resolve_body(it.f, it.depth, it.sub_count, it.bs.skip, ne, it.colour, it.accum, 1, it.stats, it.mom, it.zmin,
                 blockIdx.x - item_blk0[item]);
}
}  // namespace

// ================================================================== launchers
static inline unsigned blocks_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_view(hipStream_t s, const GpsFrame &f, const float4 *splats, const float *sh, float4 *view, uint2 *cinfo,
                 float4 *colour, uint32_t *perm) {
    if (f.N == 0) return;
    hipLaunchKernelGGL(k_view, dim3(blocks_for(f.N, 256)), dim3(256), 0, s, f, splats, sh, view, cinfo, colour,
                       perm);
}

void launch_count_scan(hipStream_t s, const GpsFrame &f, const uint2 *cinfo, uint32_t first_seed, uint32_t nens,
                       uint32_t *offsets, const BatchState &bs, uint32_t tiles, uint32_t epoch, uint64_t *stats,
                       const uint32_t *perm) {
    hipLaunchKernelGGL(k_count_scan, dim3(tiles, nens), dim3(SCAN_BLOCK), 0, s, f, cinfo, first_seed, offsets, bs,
                       tiles, epoch, stats, perm);
}

void launch_particles(hipStream_t s, const GpsFrame &f, const float4 *view, const uint32_t *offsets,
                      const BatchState &bs, uint32_t nens, uint32_t first_seed, uint64_t *depth, uint64_t sub_count,
                      const float *lut, uint64_t *stats, int grid, uint64_t *pkeys, uint32_t *pidx, int ppt,
                      const uint32_t *perm, int block) {
#define GPS_LAUNCH_BS(MK, PPT, OUT, BS)                                                                            \
    hipLaunchKernelGGL((k_particles<MK, PPT, OUT, BS>), dim3(grid), dim3(BS), 0, s, f, view, offsets, bs.totals,   \
                       nens, first_seed, depth, sub_count, lut, stats, pkeys, pidx, OUT ? nullptr : perm)
#define GPS_LAUNCH(MK, PPT, OUT)                                         \
    do {                                                                 \
        if (block == 128) GPS_LAUNCH_BS(MK, PPT, OUT, 128);              \
        else if (block == 512) GPS_LAUNCH_BS(MK, PPT, OUT, 512);         \
        else GPS_LAUNCH_BS(MK, PPT, OUT, 256);                           \
    } while (0)
#define GPS_LAUNCH_PPT(MK, OUT)          \
    do {                                 \
        if (ppt == 1) GPS_LAUNCH(MK, 1, OUT); \
        else if (ppt == 2) GPS_LAUNCH(MK, 2, OUT); \
        else if (ppt == 8) GPS_LAUNCH(MK, 8, OUT); \
        else GPS_LAUNCH(MK, 4, OUT);     \
    } while (0)
    const int mk = f.method == M_GPS ? 0 : (f.method == M_OCCUPANCY ? 2 : 1);
    if (pkeys) {
        if (mk == 0) GPS_LAUNCH_PPT(0, 1);
        else if (mk == 2) GPS_LAUNCH_PPT(2, 1);
        else GPS_LAUNCH_PPT(1, 1);
    } else {
        if (mk == 0) GPS_LAUNCH_PPT(0, 0);
        else if (mk == 2) GPS_LAUNCH_PPT(2, 0);
        else GPS_LAUNCH_PPT(1, 0);
    }
#undef GPS_LAUNCH_PPT
#undef GPS_LAUNCH
#undef GPS_LAUNCH_BS
}

void launch_resolve(hipStream_t s, const GpsFrame &f, uint64_t *depth, uint64_t sub_count, const uint32_t *skip,
                    uint32_t nens, const float4 *colour, float4 *accum, int clear, uint64_t *stats, float4 *mom,
                    float *zmin) {
    uint64_t X = (uint64_t)f.W * f.H;
    hipLaunchKernelGGL(k_resolve, dim3(blocks_for(X, 256)), dim3(256), 0, s, f, depth, sub_count, skip, nens, colour,
                       accum, clear, stats, mom, zmin);
}

void launch_features_clear(hipStream_t s, float4 *mom, float *zmin, uint64_t n) {
    if (n == 0 || !mom) return;
    hipLaunchKernelGGL(k_features_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, mom, zmin, n);
}

void launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t n, uint64_t v) {
    if (n == 0) return;
    unsigned grid = blocks_for(n, 256);
    if (grid > 8192) grid = 8192;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, s, p, n, v);
}

void launch_splat_pass(hipStream_t s, const GpsFrame &f, const float4 *view, const uint2 *cinfo, uint32_t seed,
                       uint64_t *depth, const float *lut, uint64_t *stats) {
    if (f.N == 0) return;
    hipLaunchKernelGGL(k_splat_pass, dim3(blocks_for(f.N, 64)), dim3(64), 0, s, f, view, cinfo, seed, depth, lut,
                       stats);
}

void launch_composite(hipStream_t s, const float4 *accum, uint32_t W, uint32_t H, float exposure, uint32_t gamma_on,
                      uint32_t ow, uint32_t oh, uint32_t *out) {
    hipLaunchKernelGGL(k_composite, dim3(blocks_for((uint64_t)ow * oh, 256)), dim3(256), 0, s, accum, W, H, exposure,
                       gamma_on, ow, oh, out);
}

void launch_zero_u32(hipStream_t s, uint32_t *p, uint64_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
}

void launch_stats_add(hipStream_t s, uint64_t *stats, int slot, uint64_t v) {
    hipLaunchKernelGGL(k_stats_add, dim3(1), dim3(64), 0, s, stats, slot, v);
}

void launch_stats_clear(hipStream_t s, uint64_t *stats, int first, int count) {
    hipLaunchKernelGGL(k_stats_clear, dim3(1), dim3(64), 0, s, stats, first, count);
}

void launch_copy_f4(hipStream_t s, float4 *dst, const float4 *src, uint64_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_copy_f4, dim3(blocks_for(n, 256)), dim3(256), 0, s, dst, src, n);
}

void launch_accum_clear(hipStream_t s, float4 *accum, uint64_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_accum_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, accum, n);
}

void launch_phase0(hipStream_t s, uint32_t seed, uint32_t pixels, uint32_t samples, uint64_t *buf) {
    hipLaunchKernelGGL(k_phase0, dim3(blocks_for(samples, 256)), dim3(256), 0, s, seed, pixels, samples, buf);
}

void launch_scan_u32(hipStream_t s, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *ticket, uint64_t *status,
                     uint32_t tiles, uint32_t epoch) {
    hipLaunchKernelGGL(k_scan_u32, dim3(tiles), dim3(SCAN_BLOCK), 0, s, in, out, n, ticket, status, tiles, epoch);
}

void launch_math(hipStream_t s, int fn, const float *x, const float *y, float *out, uint64_t n) {
    hipLaunchKernelGGL(k_math, dim3(blocks_for(n, 256)), dim3(256), 0, s, fn, x, y, out, n);
}

void launch_fused_clear(hipStream_t s, const FusedItem *items, uint32_t nitems, uint32_t max_x) {
    hipLaunchKernelGGL(k_fused_clear, dim3(max(1u, min(4096u, (max_x + 255u) / 256u)), nitems), dim3(256), 0, s,
                       items);
}

void launch_fused_view(hipStream_t s, const FusedItem *items, const uint32_t *blk_item, const uint32_t *item_blk0,
                       uint32_t blocks) {
    if (blocks) hipLaunchKernelGGL(k_view_fused, dim3(blocks), dim3(256), 0, s, items, blk_item, item_blk0);
}

void launch_fused_count_scan(hipStream_t s, const FusedItem *items, const uint32_t *blk_pair, uint32_t blocks,
                             uint32_t ens0, uint32_t epoch) {
    if (blocks)
        hipLaunchKernelGGL(k_count_scan_fused, dim3(blocks), dim3(SCAN_BLOCK), 0, s, items, blk_pair, ens0, epoch);
}

void launch_fused_assign(hipStream_t s, const FusedItem *items, uint32_t nitems, const uint32_t *nens, uint32_t mk,
                         uint32_t chunk, uint32_t grid, uint32_t *blk_item, uint32_t *blk_rank, uint32_t *item_nblk) {
    hipLaunchKernelGGL(k_fused_assign, dim3(1), dim3(FUSED_MAX_ITEMS), 0, s, items, nitems, nens, mk, chunk, grid,
                       blk_item, blk_rank, item_nblk);
}

void launch_fused_particles(hipStream_t s, const FusedItem *items, uint32_t mk, const uint32_t *nens, uint32_t ens0,
                            const uint32_t *blk_item, const uint32_t *blk_rank, const uint32_t *item_nblk, int grid,
                            int ppt, int block) {
#define GPS_FUSED_BS(MK, PPT, BS)                                                                             \
    hipLaunchKernelGGL((k_particles_fused<MK, PPT, BS>), dim3(grid), dim3(BS), 0, s, items, nens, ens0,      \
                       blk_item, blk_rank, item_nblk)
#define GPS_FUSED(MK, PPT)                                       \
    do {                                                         \
        if (block == 128) GPS_FUSED_BS(MK, PPT, 128);            \
        else if (block == 512) GPS_FUSED_BS(MK, PPT, 512);       \
        else GPS_FUSED_BS(MK, PPT, 256);                         \
    } while (0)
#define GPS_FUSED_PPT(MK)                         \
    do {                                          \
        if (ppt == 1) GPS_FUSED(MK, 1);           \
        else if (ppt == 2) GPS_FUSED(MK, 2);      \
        else if (ppt == 8) GPS_FUSED(MK, 8);      \
        else GPS_FUSED(MK, 4);                    \
    } while (0)
    if (mk == 0) GPS_FUSED_PPT(0);
    else if (mk == 2) GPS_FUSED_PPT(2);
    else GPS_FUSED_PPT(1);
#undef GPS_FUSED_PPT
#undef GPS_FUSED
#undef GPS_FUSED_BS
}

void launch_fused_resolve(hipStream_t s, const FusedItem *items, const uint32_t *nens, const uint32_t *blk_item,
                          const uint32_t *item_blk0, uint32_t blocks) {
    if (blocks) hipLaunchKernelGGL(k_resolve_fused, dim3(blocks), dim3(256), 0, s, items, nens, blk_item, item_blk0);
}
