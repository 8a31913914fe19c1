/*
 * CPU ORACLE -- test infrastructure only.  Never linked into the product
 * library; only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
 * load it.
 *
 * f32 restatement of the reference's GPU pipeline, sequential and in plain C:
 *   shaders/gps_splat.wgsl    (whole: prepare, particle passes, resolve,
 *                              splat-parallel path, sorted-path projection)
 *   shaders/gps_scan.wgsl     (exclusive u32 scan: a serial prefix sum)
 *   shaders/gps_sort.wgsl     (stable LSD radix sort: a stable counting sort)
 *   shaders/gps_sorted_draw.wgsl + sorted.rs:379-399 (back-to-front blend)
 *   shaders/gps_composite.wgsl (bilinear, exposure, clamp, sRGB OETF)
 *   crates/gps-render/src/gps.rs:812-925 (uniform packing, ensemble loop)
 *
 * Semantics kept from WGSL: u32 arithmetic wraps; f32 -> u32 truncates and
 * saturates; round() is half-to-even; mix(a,b,t) = a(1-t) + b t; vector
 * dot products and sums are evaluated left to right; the uniform is the top
 * 24 bits of the PCG state.  exp/log/sin/cos are the fixed algorithms of
 * f32math.h (WGSL leaves their bits to the driver).  Compile with
 * -ffp-contract=off.
 *
 * The depth and colour passes are kept as two separate passes with 32-bit
 * minima, exactly as the reference runs them (gps_splat.wgsl:276-281,
 * 526-531), so that a test comparing this oracle with a single packed 64-bit
 * minimum checks that the two are the same.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "f32math.h"

#define EXPORT __attribute__((visibility("default")))

typedef struct { float pos_opacity[4], log_scale[4], rot_wxyz[4], color[4]; } MGaussian;   /* gps.rs:29-38 */
typedef struct { double view_rot[9], view_pos[3], focal_x, focal_y, cx, cy, low_pass; } MCamera; /* oracle.rs:33-42 */
typedef struct {
    uint32_t calibration, radial_correction, centre_depth, low_pass_jitter;
    float base_k, reference_pixel_length;
} MPbvr;                                                                                   /* gps.rs:169-179 */
typedef struct {
    uint32_t spp_side;
    float density_scale;
    uint32_t max_points_per_splat;
    float near;
    float background[3];
    float exposure;
    uint32_t srgb_output, exact_radius, sh_degree, path, method;
    MPbvr pbvr;
} MParams;                                                                                 /* gps.rs:189-210 */

/* FrameUniform, gps.rs:63-76 */
typedef struct {
    float rot[3][4];
    float pos[4];
    float intr[4];
    uint32_t dims[4];
    uint32_t cfg[4];
    uint32_t cfg2[4];
    float prm[4];
    float bg[4];
    uint32_t pbvr[4];
    float pbp[4];
} MFrame;

/* stats slots of this oracle (u64 each) */
enum { S_POINTS, S_TRUNC, S_DROPPED, S_VISIBLE, S_ORPHAN, S_CAND, S_UNDER, S_SKIPPED, S_ENS, S_PROBE, S_WORDS = 16 };

#define PI 3.14159265358979f
#define DILOG_ONE 1.64493406684823f
#define EMPTY 0xFFFFFFFFu
#define METHOD_PROPORTIONAL 1u
#define METHOD_VIEW_CONDITIONED 3u
#define METHOD_OCCUPANCY 4u
#define FLAG_RADIAL 1u
#define FLAG_CENTRE 2u
#define FLAG_JITTER 4u

static inline uint32_t fbits(float x) { return om_bits(x); }

/* ------------------------------------------------------------ renderer state */
typedef struct {
    const MFrame *f;
    const MGaussian *g;
    const float *sh;
    const float *lut;
    uint32_t *depth;
    uint32_t *winner;
    uint64_t *stats;
    float (*proj)[4];       /* 6 x vec4 per splat */
    uint32_t rng;
    float spare;
    int has_spare;
} Ctx;

/* gps_splat.wgsl:57-61 */
EXPORT uint32_t mir_pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

/* gps_splat.wgsl:63-68 */
EXPORT uint32_t mir_particle_seed(uint32_t splat, uint32_t k, uint32_t frame) {
    uint32_t s = mir_pcg(splat * 747796405u + k + 1u);
    s = mir_pcg(s ^ (frame * 2654435761u + 1u));
    return s ? s : 1u;
}

/* gps_splat.wgsl:73-76 */
static inline float nextf(Ctx *c) {
    c->rng = mir_pcg(c->rng);
    return (float)(c->rng >> 8u) * (1.0f / 16777216.0f);
}

/* gps_splat.wgsl:79-87 */
static float dilog_series(float x) {
    float term = x, sum = x;
    for (uint32_t k = 2u; k < 32u; k++) {
        term = term * x;
        sum = sum + term / (float)(k * k);
    }
    return sum;
}

/* gps_splat.wgsl:89-94 */
EXPORT float mir_dilog(float x) {
    if (x >= 1.0f) return DILOG_ONE;
    if (x <= 0.0f) return 0.0f;
    if (x <= 0.5f) return dilog_series(x);
    return DILOG_ONE - om_logf(x) * om_logf(1.0f - x) - dilog_series(1.0f - x);
}

/* gps_splat.wgsl:97-113 */
EXPORT float mir_inv_dilog(float tgt, float hi_in) {
    if (tgt <= 0.0f) return 0.0f;
    float lo = 0.0f, hi = hi_in, y = hi * 0.5f;
    for (uint32_t i = 0; i < 24u; i++) {
        float f = mir_dilog(y) - tgt;
        if (f > 0.0f) hi = y; else lo = y;
        float dfdy = 1.0f;
        if (y > 1e-6f && y < 0.9999999f) dfdy = -om_logf(1.0f - y) / y;
        float ny = y - f / dfdy;
        if (!(ny > lo && ny < hi)) ny = 0.5f * (lo + hi);
        float dd = ny - y;
        if (dd < 0.0f) dd = -dd;
        if (dd <= 1e-6f * (1.0f + y)) return ny;
        y = ny;
    }
    return y;
}

/* gps_splat.wgsl:117-130 */
EXPORT float mir_corrected_radius_lut(const float *lut, float o, float u) {
    if (o <= 0.0f) return 0.0f;
    float v = -om_logf(om_max(1.0f - u, 5.96e-8f));
    float fo = (1.0f - __builtin_sqrtf(om_max(1.0f - om_min(o, 1.0f), 0.0f))) * 64.0f;
    float fv = om_min(__builtin_sqrtf(v) * (float)(128.0 / 4.123105625617661), 128.0f);
    uint32_t i0 = om_f2u(fo);
    if (i0 > 63u) i0 = 63u;
    uint32_t j0 = om_f2u(fv);
    if (j0 > 127u) j0 = 127u;
    float a = fo - (float)i0, b = fv - (float)j0;
    uint32_t k = i0 * 129u + j0;
    float r0 = om_mix(lut[k], lut[k + 1u], b);
    float r1 = om_mix(lut[k + 129u], lut[k + 130u], b);
    return om_mix(r0, r1, a);
}

/* gps_splat.wgsl:133-140 */
EXPORT float mir_corrected_radius(float o, float u) {
    if (o <= 0.0f) return 0.0f;
    float d = mir_dilog(o);
    float tgt = (1.0f - u) * d;
    float y = om_min(mir_inv_dilog(tgt, o), o);
    if (y <= 0.0f) return 40.0f;
    return __builtin_sqrtf(om_max(-2.0f * om_logf(y / o), 0.0f));
}

/* gps_splat.wgsl:142-160 */
static uint32_t poisson(Ctx *c, float lambda) {
    if (lambda <= 0.0f) return 0u;
    if (lambda < 30.0f) {
        float l = om_expf(-lambda);
        uint32_t k = 0u;
        float p = 1.0f;
        for (;;) {
            k = k + 1u;
            p = p * nextf(c);
            if (p <= l) break;
        }
        return k - 1u;
    }
    float u1 = om_max(nextf(c), 5.96e-8f);
    float u2 = nextf(c);
    float z = __builtin_sqrtf(-2.0f * om_logf(u1)) * om_cosf(2.0f * PI * u2);
    return om_f2u(om_max(om_rint(lambda + __builtin_sqrtf(lambda) * z), 0.0f));
}

/* test hook: Poisson draw from the stream seeded `seed` */
EXPORT uint32_t mir_poisson(float lambda, uint32_t seed) {
    Ctx c;
    memset(&c, 0, sizeof c);
    c.rng = seed;
    return poisson(&c, lambda);
}

typedef struct { int ok; float mx, my, a, b, c, depth; } Proj;

static inline float dot3f(const float *r, float x, float y, float z) { return r[0] * x + r[1] * y + r[2] * z; }

/* gps_splat.wgsl:165-206 */
static Proj project(const Ctx *cx, uint32_t gid) {
    const MFrame *f = cx->f;
    const MGaussian *g = &cx->g[gid];
    Proj pr;
    memset(&pr, 0, sizeof pr);
    float px = g->pos_opacity[0] - f->pos[0], py = g->pos_opacity[1] - f->pos[1], pz = g->pos_opacity[2] - f->pos[2];
    float camx = dot3f(f->rot[0], px, py, pz), camy = dot3f(f->rot[1], px, py, pz), camz = dot3f(f->rot[2], px, py, pz);
    if (camz <= om_max(f->prm[1], 1e-4f)) return pr;
    const float lim = 5.2f;                       /* 1.3 * 4.0 */
    float tx = om_clamp(camx / camz, -lim, lim) * camz;
    float ty = om_clamp(camy / camz, -lim, lim) * camz;
    float fx = f->intr[0], fy = f->intr[1];
    float iz = 1.0f / camz;
    float sx = fx * iz, sxz = -fx * tx * iz * iz;
    float sy = fy * iz, syz = -fy * ty * iz * iz;
    float m0[3], m1[3];
    for (int j = 0; j < 3; j++) {
        m0[j] = f->rot[0][j] * sx + f->rot[2][j] * sxz;
        m1[j] = f->rot[1][j] * sy + f->rot[2][j] * syz;
    }
    const float *q4 = g->rot_wxyz;
    float ql = __builtin_sqrtf(q4[0] * q4[0] + q4[1] * q4[1] + q4[2] * q4[2] + q4[3] * q4[3]);
    float w = q4[0] / ql, x = q4[1] / ql, y = q4[2] / ql, z = q4[3] / ql;
    float c0[3] = {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y)};
    float c1[3] = {2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x)};
    float c2[3] = {2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)};
    float s2x = om_expf(2.0f * g->log_scale[0]), s2y = om_expf(2.0f * g->log_scale[1]), s2z = om_expf(2.0f * g->log_scale[2]);
    float ux = dot3f(m0, c0[0], c0[1], c0[2]), uy = dot3f(m0, c1[0], c1[1], c1[2]), uz = dot3f(m0, c2[0], c2[1], c2[2]);
    float vx = dot3f(m1, c0[0], c0[1], c0[2]), vy = dot3f(m1, c1[0], c1[1], c1[2]), vz = dot3f(m1, c2[0], c2[1], c2[2]);
    float a = s2x * (ux * ux) + s2y * (uy * uy) + s2z * (uz * uz) + f->pos[3];
    float b = s2x * (ux * vx) + s2y * (uy * vy) + s2z * (uz * vz);
    float c = s2x * (vx * vx) + s2y * (vy * vy) + s2z * (vz * vz) + f->pos[3];
    float det = a * c - b * b;
    if (!(det > 0.0f)) return pr;
    pr.ok = 1;
    pr.mx = fx * camx / camz + f->intr[2];
    pr.my = fy * camy / camz + f->intr[3];
    pr.a = a;
    pr.b = b;
    pr.c = c;
    pr.depth = camz;
    return pr;
}

/* gps_splat.wgsl:210-251 */
static void splat_colour(const Ctx *cx, uint32_t id, float out[3]) {
    const MFrame *f = cx->f;
    const MGaussian *g = &cx->g[id];
    float res[3] = {g->color[0], g->color[1], g->color[2]};
    uint32_t deg = f->pbvr[3];
    if (deg == 0u) {
        for (int k = 0; k < 3; k++) out[k] = om_max(res[k], 0.0f);
        return;
    }
    float px = g->pos_opacity[0] - f->pos[0], py = g->pos_opacity[1] - f->pos[1], pz = g->pos_opacity[2] - f->pos[2];
    float len = __builtin_sqrtf(px * px + py * py + pz * pz);
    float x = px / len, y = py / len, z = pz / len;
    uint32_t r = f->cfg2[1];
    uint32_t o = id * 3u * r;
    const float *sh = cx->sh;
#define S(k, ch) sh[o + (ch) * r + (k)]
    for (int ch = 0; ch < 3; ch++) {
        float v = res[ch];
        v = v - 0.4886025119029199f * y * S(0, ch) + 0.4886025119029199f * z * S(1, ch) -
            0.4886025119029199f * x * S(2, ch);
        if (deg >= 2u) {
            float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            v = v + 1.0925484305920792f * xy * S(3, ch) - 1.0925484305920792f * yz * S(4, ch) +
                0.31539156525252005f * (2.0f * zz - xx - yy) * S(5, ch) - 1.0925484305920792f * xz * S(6, ch) +
                0.5462742152960396f * (xx - yy) * S(7, ch);
            if (deg >= 3u) {
                v = v - 0.5900435899266435f * y * (3.0f * xx - yy) * S(8, ch) + 2.890611442640554f * xy * z * S(9, ch) -
                    0.4570457994644658f * y * (4.0f * zz - xx - yy) * S(10, ch) +
                    0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * S(11, ch) -
                    0.4570457994644658f * x * (4.0f * zz - xx - yy) * S(12, ch) +
                    1.445305721320277f * z * (xx - yy) * S(13, ch) -
                    0.5900435899266435f * x * (xx - 3.0f * yy) * S(14, ch);
            }
        }
        out[ch] = om_max(v, 0.0f);
    }
#undef S
}

/* gps_splat.wgsl:47-50 (64-bit counters) */
static inline void stat_add(Ctx *c, int k, uint64_t n) { c->stats[k] += n; }

/* atomic u32 minimum (the particle loop may run on several threads; the result is order independent) */
static inline void atomic_min_u32(uint32_t *p, uint32_t v) {
    uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

/* the two 32-bit passes of gps_splat.wgsl:276-281 / 526-531 */
static inline void write_subpixel(Ctx *c, uint32_t mode, uint32_t idx, uint32_t key, uint32_t id) {
    if (mode == 0u) {
        atomic_min_u32(&c->depth[idx], key);
    } else if (c->depth[idx] == key) {
        atomic_min_u32(&c->winner[idx], id);
    }
}

/* gps_splat.wgsl:259-282 */
static void emit_particle(Ctx *c, uint32_t gid, uint32_t mode, float mx, float my, float depth_z, float o, float l00,
                          float l10, float l11) {
    const MFrame *f = c->f;
    uint32_t spp_side = f->dims[2], spp = spp_side * spp_side, W = f->dims[0], H = f->dims[1];
    float r;
    if (f->cfg2[2] == 0u) {
        float u = nextf(c);
        r = mir_corrected_radius_lut(c->lut, o, u);
    } else {
        float u = nextf(c);
        r = mir_corrected_radius(o, u);
    }
    float th = 2.0f * PI * nextf(c);
    float wx = r * om_cosf(th), wy = r * om_sinf(th);
    float pxf = mx + l00 * wx, pyf = my + (l10 * wx + l11 * wy);
    if (pxf < 0.0f || pxf >= (float)W || pyf < 0.0f || pyf >= (float)H) return;
    uint32_t px = om_f2u(pxf), py = om_f2u(pyf);
    uint32_t sx = om_f2u((pxf - (float)px) * (float)spp_side);
    uint32_t sy = om_f2u((pyf - (float)py) * (float)spp_side);
    if (sx > spp_side - 1u) sx = spp_side - 1u;
    if (sy > spp_side - 1u) sy = spp_side - 1u;
    uint32_t idx = (py * W + px) * spp + sy * spp_side + sx;
    uint32_t key = fbits(depth_z) | 0x80000000u;
    write_subpixel(c, mode, idx, key, gid);
}

/* gps_splat.wgsl:288-293 */
static int offscreen(const Ctx *c, const Proj *pr) {
    float ex = 5.9f * __builtin_sqrtf(pr->a), ey = 5.9f * __builtin_sqrtf(pr->c);
    return pr->mx + ex < 0.0f || pr->mx - ex >= (float)c->f->dims[0] || pr->my + ey < 0.0f ||
           pr->my - ey >= (float)c->f->dims[1];
}

/* gps_splat.wgsl:296-300 */
static float expected_count(const Ctx *c, const Proj *pr, float o) {
    uint32_t spp = c->f->dims[2] * c->f->dims[2];
    float det = pr->a * pr->c - pr->b * pr->b;
    return c->f->prm[0] * (float)spp * 2.0f * PI * __builtin_sqrtf(det) * mir_dilog(o);
}

/* gps_splat.wgsl:304-333 */
static void splat_pass(Ctx *c, uint32_t gid, uint32_t mode) {
    Proj pr = project(c, gid);
    if (!pr.ok) return;
    float o = om_clamp(c->g[gid].pos_opacity[3], 0.0f, 1.0f);
    if (o <= 0.0f) return;
    if (offscreen(c, &pr)) return;
    float en = expected_count(c, &pr, o);
    if (!(en > 0.0f)) return;
    c->rng = mir_particle_seed(gid, 0u, c->f->cfg[0]);
    uint32_t n = poisson(c, en);
    if (mode == 0u) {
        stat_add(c, S_POINTS, n);
        stat_add(c, S_VISIBLE, 1);
    }
    uint32_t cap = c->f->cfg[1];
    if (n > cap) {
        if (mode == 0u) {
            stat_add(c, S_TRUNC, 1);
            stat_add(c, S_DROPPED, n - cap);
        }
        n = cap;
    }
    float l00 = __builtin_sqrtf(om_max(pr.a, 1e-30f));
    float l10 = pr.b / l00;
    float l11 = __builtin_sqrtf(om_max(pr.c - l10 * l10, 1e-30f));
    for (uint32_t i = 0; i < n; i++) emit_particle(c, gid, mode, pr.mx, pr.my, pr.depth, o, l00, l10, l11);
}

/* gps_splat.wgsl:402-414 */
static float normal(Ctx *c) {
    if (c->has_spare) {
        c->has_spare = 0;
        return c->spare;
    }
    float u1 = om_max(nextf(c), 5.96e-8f);
    float u2 = nextf(c);
    float r = __builtin_sqrtf(-2.0f * om_logf(u1));
    float th = 2.0f * PI * u2;
    c->spare = r * om_sinf(th);
    c->has_spare = 1;
    return r * om_cosf(th);
}

/* gps_splat.wgsl:417-421 */
static float extinction_term(float o) {
    float c = om_min(o, 1.0f - 1e-6f);
    if (c < 1e-3f) return c * (1.0f + c * (0.5f + c * 0.33333334f));
    return -om_logf(1.0f - c);
}

/* gps_splat.wgsl:423-427 */
static float pixel_density_scale(const MFrame *f, float depth) {
    if (!(depth > om_max(f->prm[1], 0.0f))) return 0.0f;
    float ratio = f->pbp[1] / depth;
    return om_min(ratio * ratio * f->intr[0] * f->intr[1], 1.0e12f);
}

/* gps_splat.wgsl:430-438 */
static float extinction_over_linear(float x) {
    if (x < 1e-4f) return 1.0f + 0.5f * x;
    return -om_logf(1.0f - om_min(x, 1.0f - 1e-6f)) / x;
}
EXPORT float mir_radial_keep(float o, float r2) {
    float oo = om_min(o, 1.0f - 1e-6f);
    float y = oo * om_expf(-0.5f * om_max(r2, 0.0f));
    return om_clamp(extinction_over_linear(y) / extinction_over_linear(oo), 0.0f, 1.0f);
}

/* gps_splat.wgsl:441-460 */
static float pbvr_lambda(const Ctx *c, const Proj *pr, float o) {
    const MFrame *f = c->f;
    uint32_t method = f->pbvr[0];
    float g = extinction_term(o);
    if (method == METHOD_PROPORTIONAL) g = om_min(o, 1.0f - 1e-6f);
    float scale = f->pbp[0];
    if (method != METHOD_VIEW_CONDITIONED) {
        uint32_t level = f->pbvr[1];
        if (level == 1u) {
            scale = scale * pixel_density_scale(f, f->pbp[2]);
        } else if (level == 2u) {
            scale = scale * pixel_density_scale(f, pr->depth);
        } else if (level == 3u) {
            float spp = (float)(f->dims[2] * f->dims[2]);
            scale = spp * 2.0f * PI * __builtin_sqrtf(pr->a * pr->c - pr->b * pr->b);
        }
    }
    float n = f->prm[0] * g * scale;
    if (!(n > 0.0f) || n > 3.0e38f) return 0.0f;
    return n;
}

/* gps_splat.wgsl:464-485 */
static void chol_of(const Ctx *c, uint32_t gid, float ca[3], float cb[3]) {
    const MGaussian *g = &c->g[gid];
    const float *q4 = g->rot_wxyz;
    float ql = __builtin_sqrtf(q4[0] * q4[0] + q4[1] * q4[1] + q4[2] * q4[2] + q4[3] * q4[3]);
    float w = q4[0] / ql, x = q4[1] / ql, y = q4[2] / ql, z = q4[3] / ql;
    float c0[3] = {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y)};
    float c1[3] = {2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x)};
    float c2[3] = {2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y)};
    float sx = om_expf(2.0f * g->log_scale[0]), sy = om_expf(2.0f * g->log_scale[1]), sz = om_expf(2.0f * g->log_scale[2]);
    float a00 = sx * c0[0] * c0[0] + sy * c1[0] * c1[0] + sz * c2[0] * c2[0];
    float a10 = sx * c0[1] * c0[0] + sy * c1[1] * c1[0] + sz * c2[1] * c2[0];
    float a20 = sx * c0[2] * c0[0] + sy * c1[2] * c1[0] + sz * c2[2] * c2[0];
    float a11 = sx * c0[1] * c0[1] + sy * c1[1] * c1[1] + sz * c2[1] * c2[1];
    float a21 = sx * c0[2] * c0[1] + sy * c1[2] * c1[1] + sz * c2[2] * c2[1];
    float a22 = sx * c0[2] * c0[2] + sy * c1[2] * c1[2] + sz * c2[2] * c2[2];
    float l00 = __builtin_sqrtf(om_max(a00, 1e-30f));
    float l10 = a10 / l00;
    float l20 = a20 / l00;
    float l11 = __builtin_sqrtf(om_max(a11 - l10 * l10, 1e-30f));
    float l21 = (a21 - l20 * l10) / l11;
    float l22 = __builtin_sqrtf(om_max(a22 - l20 * l20 - l21 * l21, 1e-30f));
    ca[0] = l00; ca[1] = l10; ca[2] = l11;
    cb[0] = l20; cb[1] = l21; cb[2] = l22;
}

/* gps_splat.wgsl:490-532; a,b,c,d = proj rows 0..3, ch = rows 4, 5 */
static void emit_particle_pbvr(Ctx *c, uint32_t gid, uint32_t mode, const float *a, const float *b, const float *cc,
                               const float *d, const float *ca, const float *cb) {
    const MFrame *f = c->f;
    c->has_spare = 0;
    float uz = normal(c);
    float uy = normal(c);
    float ux = normal(c);
    if (f->pbvr[0] == METHOD_VIEW_CONDITIONED) {
        if (nextf(c) >= b[3]) return;
    }
    float wx = cc[0] + ca[0] * ux;
    float wy = cc[1] + (ca[1] * ux + ca[2] * uy);
    float wz = cc[2] + (cb[0] * ux + cb[1] * uy + cb[2] * uz);
    float px = wx - f->pos[0], py = wy - f->pos[1], pz = wz - f->pos[2];
    float camx = dot3f(f->rot[0], px, py, pz), camy = dot3f(f->rot[1], px, py, pz), camz = dot3f(f->rot[2], px, py, pz);
    if (camz <= f->prm[1]) return;
    float pixx = f->intr[0] * camx / camz + f->intr[2];
    float pixy = f->intr[1] * camy / camz + f->intr[3];
    uint32_t flags = f->pbvr[2];
    if (flags & FLAG_JITTER) {
        float jy = normal(c);
        float jx = normal(c);
        float s = __builtin_sqrtf(om_max(f->pos[3], 0.0f));
        pixx = pixx + s * jx;
        pixy = pixy + s * jy;
    }
    if (flags & FLAG_RADIAL) {
        float dvx = pixx - a[0], dvy = pixy - a[1];
        float r2 = cc[3] * dvx * dvx + 2.0f * d[0] * dvx * dvy + d[1] * dvy * dvy;
        if (nextf(c) >= mir_radial_keep(a[3], r2)) return;
    }
    float depth_z = camz;
    if (flags & FLAG_CENTRE) depth_z = a[2];
    if (mode == 0u) stat_add(c, S_POINTS, 1);
    uint32_t W = f->dims[0], H = f->dims[1];
    if (pixx < 0.0f || pixx >= (float)W || pixy < 0.0f || pixy >= (float)H) return;
    uint32_t spp_side = f->dims[2];
    uint32_t ipx = om_f2u(pixx), ipy = om_f2u(pixy);
    uint32_t sx = om_f2u((pixx - (float)ipx) * (float)spp_side);
    uint32_t sy = om_f2u((pixy - (float)ipy) * (float)spp_side);
    if (sx > spp_side - 1u) sx = spp_side - 1u;
    if (sy > spp_side - 1u) sy = spp_side - 1u;
    uint32_t idx = (ipy * W + ipx) * spp_side * spp_side + sy * spp_side + sx;
    uint32_t key = fbits(depth_z) | 0x80000000u;
    write_subpixel(c, mode, idx, key, gid);
}

/* gps_splat.wgsl:535-550 */
static uint32_t prepare_gps(Ctx *c, uint32_t i, const Proj *pr, float o) {
    if (offscreen(c, pr)) return 0u;
    float en = expected_count(c, pr, o);
    if (!(en > 0.0f)) return 0u;
    c->rng = mir_particle_seed(i, 0u, c->f->cfg[0]);
    uint32_t n = poisson(c, en);
    stat_add(c, S_POINTS, n);
    stat_add(c, S_VISIBLE, 1);
    uint32_t cap = c->f->cfg[1];
    if (n > cap) {
        stat_add(c, S_TRUNC, 1);
        stat_add(c, S_DROPPED, n - cap);
        n = cap;
    }
    return n;
}

/* gps_splat.wgsl:552-612 */
static uint32_t prepare_pbvr(Ctx *c, uint32_t i, const Proj *pr, float o) {
    const MFrame *f = c->f;
    float(*P)[4] = &c->proj[(size_t)i * 6u];
    if (f->pbvr[0] == METHOD_OCCUPANCY) {
        float density = om_max(f->prm[0], 1e-8f);
        float amin = om_max(1.0f - om_expf(om_logf(1.0f - 1.0f / 255.0f) / density), 1e-8f);
        if (o <= amin) return 0u;
        float radius = __builtin_sqrtf(2.0f * om_logf(o / amin));
        float ex = radius * __builtin_sqrtf(pr->a), ey = radius * __builtin_sqrtf(pr->c);
        float Wf = (float)f->dims[0], Hf = (float)f->dims[1];
        float lox = om_clamp(__builtin_floorf(pr->mx - ex), 0.0f, Wf), loy = om_clamp(__builtin_floorf(pr->my - ey), 0.0f, Hf);
        float hix = om_clamp(__builtin_ceilf(pr->mx + ex), 0.0f, Wf), hiy = om_clamp(__builtin_ceilf(pr->my + ey), 0.0f, Hf);
        float sx = om_max(hix - lox, 0.0f), sy = om_max(hiy - loy, 0.0f);
        uint32_t count = om_f2u(sx) * om_f2u(sy) * f->dims[2] * f->dims[2];
        if (count == 0u) return 0u;
        float invdet = 1.0f / om_max(pr->a * pr->c - pr->b * pr->b, 1e-30f);
        P[1][0] = lox; P[1][1] = loy; P[1][2] = sx; P[1][3] = sy;
        float peak_tau = -om_logf(1.0f - om_min(o, 1.0f - 1e-6f));
        int sparse = density * peak_tau < 1.0f;
        uint32_t trials = count;
        if (sparse) {
            c->rng = mir_particle_seed(i, 0u, f->cfg[0]);
            trials = poisson(c, (float)count * density * peak_tau);
        }
        P[2][0] = pr->c * invdet; P[2][1] = -pr->b * invdet; P[2][2] = pr->a * invdet; P[2][3] = peak_tau;
        P[3][3] = sparse ? 1.0f : 0.0f;
        stat_add(c, S_CAND, trials);
        stat_add(c, S_VISIBLE, 1);
        return trials;
    }
    float lambda = pbvr_lambda(c, pr, o);
    if (!(lambda > 0.0f)) return 0u;
    float det = pr->a * pr->c - pr->b * pr->b;
    float keep = 1.0f;
    if (f->pbvr[0] == METHOD_VIEW_CONDITIONED) {
        float spp = (float)(f->dims[2] * f->dims[2]);
        float target = f->prm[0] * spp * 2.0f * PI * __builtin_sqrtf(det) * mir_dilog(o);
        if (lambda > 0.5f) keep = om_clamp(target / lambda, 0.0f, 1.0f); else keep = 0.0f;
        if (target > lambda) stat_add(c, S_UNDER, 1);
    }
    c->rng = mir_particle_seed(i, 0u, f->cfg[0]);
    uint32_t n = poisson(c, lambda);
    stat_add(c, S_CAND, n);
    stat_add(c, S_VISIBLE, 1);
    uint32_t cap = f->cfg[1];
    if (n > cap) {
        stat_add(c, S_TRUNC, 1);
        stat_add(c, S_DROPPED, n - cap);
        n = cap;
    }
    float idet = 1.0f / om_max(det, 1e-30f);
    const MGaussian *g = &c->g[i];
    P[1][3] = keep;
    P[2][0] = g->pos_opacity[0]; P[2][1] = g->pos_opacity[1]; P[2][2] = g->pos_opacity[2]; P[2][3] = pr->c * idet;
    P[3][0] = -pr->b * idet; P[3][1] = pr->a * idet; P[3][2] = 0.0f; P[3][3] = 0.0f;
    float ca[3], cb[3];
    chol_of(c, i, ca, cb);
    P[4][0] = ca[0]; P[4][1] = ca[1]; P[4][2] = ca[2]; P[4][3] = 0.0f;
    P[5][0] = cb[0]; P[5][1] = cb[1]; P[5][2] = cb[2]; P[5][3] = 0.0f;
    return n;
}

/* gps_splat.wgsl:616-642 (without the overflow bookkeeping, done by the caller) */
static uint32_t prepare(Ctx *c, uint32_t i) {
    uint32_t cnt = 0u;
    Proj pr = project(c, i);
    float o = om_clamp(c->g[i].pos_opacity[3], 0.0f, 1.0f);
    if (pr.ok && o > 0.0f) {
        float l00 = __builtin_sqrtf(om_max(pr.a, 1e-30f));
        float l10 = pr.b / l00;
        float l11 = __builtin_sqrtf(om_max(pr.c - l10 * l10, 1e-30f));
        float(*P)[4] = &c->proj[(size_t)i * 6u];
        P[0][0] = pr.mx; P[0][1] = pr.my; P[0][2] = pr.depth; P[0][3] = o;
        P[1][0] = l00; P[1][1] = l10; P[1][2] = l11; P[1][3] = 1.0f;
        cnt = c->f->pbvr[0] == 0u ? prepare_gps(c, i, &pr, o) : prepare_pbvr(c, i, &pr, o);
    }
    return cnt;
}

/* gps_splat.wgsl:677-714 for particle k of splat i */
static void particle(Ctx *c, uint32_t i, uint32_t k, uint32_t mode) {
    const MFrame *f = c->f;
    const float(*P)[4] = (const float(*)[4]) & c->proj[(size_t)i * 6u];
    const float *a = P[0], *b = P[1];
    c->rng = mir_particle_seed(i, k + 1u, f->cfg[0]);
    if (f->pbvr[0] == METHOD_OCCUPANCY) {
        uint32_t side = f->dims[2], samples = side * side;
        int sparse = P[3][3] > 0.0f;
        uint32_t cell = k;
        if (sparse) {
            uint32_t a1 = om_f2u(nextf(c) * b[2] * b[3] * (float)samples);
            uint32_t a2 = om_f2u(b[2]) * om_f2u(b[3]) * samples - 1u;
            cell = a1 < a2 ? a1 : a2;
        }
        uint32_t pixel = cell / samples, sub = cell % samples;
        uint32_t bz = om_f2u(b[2]);
        uint32_t xyx = om_f2u(b[0]) + pixel % bz, xyy = om_f2u(b[1]) + pixel / bz;
        float n1 = nextf(c);
        float n2 = nextf(c);
        float uvx = ((float)(sub % side) + n1) / (float)side;
        float uvy = ((float)(sub / side) + n2) / (float)side;
        float dx = (float)xyx + uvx - a[0], dy = (float)xyy + uvy - a[1];
        const float *inv = P[2];
        float r2 = om_max(inv[0] * dx * dx + 2.0f * inv[1] * dx * dy + inv[2] * dy * dy, 0.0f);
        float alpha = om_min(a[3] * om_expf(-0.5f * r2), 1.0f - 1e-6f);
        float coverage = 1.0f - om_expf(f->prm[0] * om_logf(1.0f - alpha));
        if (sparse) coverage = -om_logf(1.0f - alpha) / inv[3];
        if (nextf(c) >= coverage) return;
        uint32_t idx = (xyy * f->dims[0] + xyx) * samples + sub;
        uint32_t key = fbits(a[2]) | 0x80000000u;
        if (mode == 0u) stat_add(c, S_POINTS, 1);
        write_subpixel(c, mode, idx, key, i);
    } else if (f->pbvr[0] == 0u) {
        emit_particle(c, i, mode, a[0], a[1], a[2], a[3], b[0], b[1], b[2]);
    } else {
        emit_particle_pbvr(c, i, mode, a, b, P[2], P[3], P[4], P[5]);
    }
}

/* Threads of the particle loop (OpenMP); 1 runs it on the calling thread only. */
EXPORT void mir_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* Denoiser feature outputs of the next mir_render_ensembles calls (NULL: off).  Test hook, not thread safe. */
static float *g_feat_mom, *g_feat_zmin;
EXPORT void mir_set_features(float *moments, float *zmin) {
    g_feat_mom = moments;
    g_feat_zmin = zmin;
}

/* gps.rs:812-856: one FrameUniform */
static void pack_frame(MFrame *fr, const MGaussian *g, uint32_t n, uint32_t sh_stride, uint32_t sh_max_degree,
                       const MCamera *cam, const MParams *p, uint32_t W, uint32_t H, uint32_t spp_side, uint32_t seed) {
    memset(fr, 0, sizeof *fr);
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < 3; k++) fr->rot[r][k] = (float)cam->view_rot[3 * r + k];
        fr->rot[r][3] = 0.0f;
    }
    for (int k = 0; k < 3; k++) fr->pos[k] = (float)cam->view_pos[k];
    fr->pos[3] = (float)cam->low_pass;
    double cx = cam->cx, cy = cam->cy;
    if (cx == 0.0 && cy == 0.0) { cx = 0.5 * (double)W; cy = 0.5 * (double)H; }
    fr->intr[0] = (float)cam->focal_x;
    fr->intr[1] = (float)cam->focal_y;
    fr->intr[2] = (float)cx;
    fr->intr[3] = (float)cy;
    fr->dims[0] = W; fr->dims[1] = H; fr->dims[2] = spp_side; fr->dims[3] = n;
    fr->cfg[0] = seed;
    fr->cfg[1] = p->max_points_per_splat > 1u ? p->max_points_per_splat : 1u;
    fr->cfg2[1] = sh_stride;
    fr->cfg2[2] = p->exact_radius ? 1u : 0u;
    fr->prm[0] = p->density_scale;
    fr->prm[1] = p->near;
    for (int k = 0; k < 3; k++) fr->bg[k] = p->background[k];
    uint32_t flags = 0;
    if (p->pbvr.radial_correction) flags |= FLAG_RADIAL;
    if (p->pbvr.centre_depth) flags |= FLAG_CENTRE;
    if (p->pbvr.low_pass_jitter) flags |= FLAG_JITTER;
    fr->pbvr[0] = p->method;
    fr->pbvr[1] = p->pbvr.calibration;
    fr->pbvr[2] = flags;
    fr->pbvr[3] = p->sh_degree < sh_max_degree ? p->sh_degree : sh_max_degree;
    /* centroid of the f32 positions in f64 (gps.rs:731-735), object depth (gps.rs:833) */
    double c[3] = {0, 0, 0};
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) c[k] = c[k] + (double)g[i].pos_opacity[k];
    if (n)
        for (int k = 0; k < 3; k++) c[k] = c[k] / (double)n;
    double d[3] = {c[0] - cam->view_pos[0], c[1] - cam->view_pos[1], c[2] - cam->view_pos[2]};
    double od = cam->view_rot[6] * d[0] + cam->view_rot[7] * d[1] + cam->view_rot[8] * d[2];
    fr->pbp[0] = p->pbvr.base_k;
    fr->pbp[1] = p->pbvr.reference_pixel_length;
    fr->pbp[2] = (float)od;
}

/*
 * gps.rs:861-925 + the per-ensemble pass list.  Returns 0, or -1 on a
 * parameter error (the reference's Err).  `accum` (W*H*4) and `stats`
 * (16 u64) accumulate.  Optional outputs describe the LAST ensemble:
 * counts_out/offsets_out (n u32), depth_id_out (W*H*spp u64 = key<<32|id,
 * all ones when empty), total_out, skipped_out.
 */
EXPORT int mir_render_ensembles(const MGaussian *g, uint32_t n, const float *sh_rest, uint32_t sh_max_degree,
                                const float *lut, const MCamera *cam, const MParams *p, uint32_t W, uint32_t H,
                                uint32_t spp_side, uint32_t first_seed, uint32_t count, float *accum, uint64_t *stats,
                                uint32_t *counts_out, uint32_t *offsets_out, uint64_t *depth_id_out,
                                uint32_t *total_out, uint32_t *skipped_out) {
    if (count == 0) return 0;
    uint32_t ps = p->spp_side < 1 ? 1 : (p->spp_side > 4 ? 4 : p->spp_side);
    if (ps != spp_side) return -1;
    uint32_t spp = spp_side * spp_side;
    size_t S = (size_t)W * H * spp, X = (size_t)W * H;
    uint32_t sh_stride = sh_max_degree ? (sh_max_degree + 1) * (sh_max_degree + 1) - 1 : 0;
    Ctx c;
    memset(&c, 0, sizeof c);
    c.g = g;
    c.sh = sh_rest;
    c.lut = lut;
    c.stats = stats;
    c.depth = (uint32_t *)malloc(S * 4);
    c.winner = (uint32_t *)malloc(S * 4);
    c.proj = (float(*)[4])calloc((size_t)(n ? n : 1) * 6u, 16);
    uint32_t *counts = (uint32_t *)calloc(n ? n : 1, 4);
    uint32_t *offs = (uint32_t *)calloc(n ? n : 1, 4);
    int use_particles = (p->path == 1u || p->method != 0u) && n > 0;
    MFrame fr;
    for (uint32_t e = 0; e < count; e++) {
        pack_frame(&fr, g, n, sh_stride, sh_max_degree, cam, p, W, H, spp_side, first_seed + e);
        c.f = &fr;
        /* clear_buffers, gps_splat.wgsl:335-347 */
        stats[S_ENS] += 1;
        for (size_t i = 0; i < S; i++) { c.depth[i] = EMPTY; c.winner[i] = EMPTY; }
        uint32_t running = 0u, overflow = 0u, total = 0u;
        if (use_particles) {
            for (uint32_t i = 0; i < n; i++) {
                uint32_t cnt = prepare(&c, i);
                if (cnt > 0u) {
                    uint32_t before = running;
                    running += cnt;
                    if (before > 0xffffffffu - cnt) overflow = 1u;
                }
                counts[i] = cnt;
            }
            /* gps_scan.wgsl: exclusive u32 prefix sum */
            uint32_t acc = 0u;
            for (uint32_t i = 0; i < n; i++) { offs[i] = acc; acc += counts[i]; }
            /* finalize_dispatch, gps_splat.wgsl:647-663 */
            total = offs[n - 1] + counts[n - 1];
            if (overflow) {
                total = 0u;
                stats[S_SKIPPED] += 1;
            }
            /* every particle is seeded by (splat, k, frame) alone, so splats can be split over threads; each
               thread keeps its own counters, added at the end (the minima are atomic) */
            for (uint32_t mode = 0; mode < 2u; mode++) {
                if (total == 0u) break;
#pragma omp parallel
                {
                    Ctx lc = c;
                    uint64_t ls[16] = {0};
                    lc.stats = ls;
#pragma omp for schedule(dynamic, 16)
                    for (int64_t si = 0; si < (int64_t)n; si++)
                        for (uint32_t k = 0; k < counts[si]; k++) particle(&lc, (uint32_t)si, k, mode);
                    for (int q = 0; q < 16; q++)
                        if (ls[q]) __atomic_fetch_add(&stats[q], ls[q], __ATOMIC_RELAXED);
                }
            }
        } else if (n > 0) {
            for (uint32_t mode = 0; mode < 2u; mode++)
                for (uint32_t i = 0; i < n; i++) splat_pass(&c, i, mode);
        }
        /* resolve, gps_splat.wgsl:361-386 */
        if (!overflow) {
            for (size_t i = 0; i < X; i++) {
                float sx = 0.0f, sy = 0.0f, sz = 0.0f;
                for (uint32_t k = 0; k < spp; k++) {
                    uint32_t d = c.depth[i * spp + k];
                    if (d == EMPTY) {
                        sx = sx + fr.bg[0]; sy = sy + fr.bg[1]; sz = sz + fr.bg[2];
                    } else {
                        uint32_t id = c.winner[i * spp + k];
                        if (id == EMPTY) {
                            stats[S_ORPHAN] += 1;
                            sx = sx + fr.bg[0]; sy = sy + fr.bg[1]; sz = sz + fr.bg[2];
                        } else {
                            float col[3];
                            splat_colour(&c, id, col);
                            sx = sx + col[0]; sy = sy + col[1]; sz = sz + col[2];
                        }
                    }
                }
                float fs = (float)spp;
                float vx = sx / fs, vy = sy / fs, vz = sz / fs;
                accum[4 * i + 0] = accum[4 * i + 0] + vx;
                accum[4 * i + 1] = accum[4 * i + 1] + vy;
                accum[4 * i + 2] = accum[4 * i + 2] + vz;
                accum[4 * i + 3] = accum[4 * i + 3] + 1.0f;
                if (g_feat_mom) {
                    /* denoiser features: squared ensemble values, covered subpixels, nearest depth */
                    uint32_t hits = 0, zb = fbits(g_feat_zmin[i]);
                    for (uint32_t k = 0; k < spp; k++) {
                        uint32_t d = c.depth[i * spp + k];
                        if (d != EMPTY) {
                            hits++;
                            if ((d & 0x7FFFFFFFu) < zb) zb = d & 0x7FFFFFFFu;
                        }
                    }
                    float *m = &g_feat_mom[4 * i];
                    m[0] = m[0] + vx * vx;
                    m[1] = m[1] + vy * vy;
                    m[2] = m[2] + vz * vz;
                    m[3] = m[3] + (float)hits;
                    memcpy(&g_feat_zmin[i], &zb, 4);
                }
            }
        }
        if (e + 1 == count) {
            if (counts_out) memcpy(counts_out, counts, (size_t)n * 4);
            if (offsets_out) memcpy(offsets_out, offs, (size_t)n * 4);
            if (total_out) *total_out = total;
            if (skipped_out) *skipped_out = overflow;
            if (depth_id_out)
                for (size_t i = 0; i < S; i++)
                    depth_id_out[i] = c.depth[i] == EMPTY ? ~0ull : ((uint64_t)c.depth[i] << 32) | c.winner[i];
        }
    }
    free(c.depth); free(c.winner); free(c.proj); free(counts); free(offs);
    return 0;
}

/* Per-splat view colour (splat_colour) for every splat: test hook. */
EXPORT void mir_splat_colours(const MGaussian *g, uint32_t n, const float *sh_rest, uint32_t sh_max_degree,
                              const MCamera *cam, const MParams *p, uint32_t W, uint32_t H, float *out /* n*3 */) {
    MFrame fr;
    uint32_t sh_stride = sh_max_degree ? (sh_max_degree + 1) * (sh_max_degree + 1) - 1 : 0;
    pack_frame(&fr, g, n, sh_stride, sh_max_degree, cam, p, W, H, p->spp_side, 0);
    Ctx c;
    memset(&c, 0, sizeof c);
    c.f = &fr;
    c.g = g;
    c.sh = sh_rest;
    for (uint32_t i = 0; i < n; i++) splat_colour(&c, i, &out[3 * i]);
}

/* ------------------------------------------------------------------ sorted path
 * sp_project (gps_splat.wgsl:736-777), stable LSD radix sort of the 24-bit keys
 * (gps_sort.wgsl + sorted.rs:221-249: a stable sort), back-to-front blend of
 * gps_sorted_draw.wgsl's quads evaluated at pixel centres in f32.
 */
EXPORT int mir_render_sorted(const MGaussian *g, uint32_t n, const float *sh_rest, uint32_t sh_max_degree,
                             const MCamera *cam, const MParams *p_in, uint32_t W, uint32_t H, uint32_t spp_side,
                             float *accum, uint64_t *stats, uint32_t *keys_out, uint32_t *vals_out, uint32_t *visible_out) {
    if (n == 0) return -1;
    MParams p = *p_in;
    p.method = 0u;
    MFrame fr;
    uint32_t sh_stride = sh_max_degree ? (sh_max_degree + 1) * (sh_max_degree + 1) - 1 : 0;
    pack_frame(&fr, g, n, sh_stride, sh_max_degree, cam, &p, W, H, spp_side, 0);
    Ctx c;
    memset(&c, 0, sizeof c);
    c.f = &fr;
    c.g = g;
    c.sh = sh_rest;
    uint32_t *keys = (uint32_t *)malloc((size_t)n * 4), *vals = (uint32_t *)malloc((size_t)n * 4);
    float(*sp)[4] = (float(*)[4])calloc((size_t)n * 3, 16);
    uint32_t visible = 0;
    stats[S_PROBE] = 0;
    for (uint32_t i = 0; i < n; i++) {
        keys[i] = 0xFFFFFFu;
        vals[i] = i;
        Proj pr = project(&c, i);
        if (!pr.ok) continue;
        float o = om_clamp(g[i].pos_opacity[3], 0.0f, 1.0f);
        if (o <= 0.0f || offscreen(&c, &pr)) continue;
        float en = expected_count(&c, &pr, o);
        if (en > 0.0f) stats[S_PROBE] += om_f2u(om_min(en, 4.0e9f));
        const float min_alpha = 0.00392156862745098f;
        if (o < min_alpha) continue;
        float k = __builtin_sqrtf(2.0f * om_logf(o / min_alpha));
        float ex = k * __builtin_sqrtf(pr.a), ey = k * __builtin_sqrtf(pr.c);
        if (pr.mx + ex < 0.0f || pr.mx - ex >= (float)W || pr.my + ey < 0.0f || pr.my - ey >= (float)H) continue;
        float tr2 = 0.5f * (pr.a + pr.c);
        float disc = __builtin_sqrtf(om_max(tr2 * tr2 - (pr.a * pr.c - pr.b * pr.b), 0.0f));
        float l1 = tr2 + disc, l2 = om_max(tr2 - disc, 1e-12f);
        float v1x = 1.0f, v1y = 0.0f;
        float absb = pr.b < 0.0f ? -pr.b : pr.b;
        if (absb > 1e-12f) {
            float ax = l1 - pr.c, ay = pr.b;
            float len = __builtin_sqrtf(ax * ax + ay * ay);
            v1x = ax / len;
            v1y = ay / len;
        } else if (pr.a < pr.c) {
            v1x = 0.0f;
            v1y = 1.0f;
        }
        float v2x = -v1y, v2y = v1x;
        float col[3];
        splat_colour(&c, i, col);
        float s1 = __builtin_sqrtf(l1), s2 = __builtin_sqrtf(l2);
        sp[3 * i][0] = pr.mx; sp[3 * i][1] = pr.my; sp[3 * i][2] = o; sp[3 * i][3] = k;
        sp[3 * i + 1][0] = v1x * s1; sp[3 * i + 1][1] = v1y * s1; sp[3 * i + 1][2] = v2x * s2; sp[3 * i + 1][3] = v2y * s2;
        sp[3 * i + 2][0] = col[0]; sp[3 * i + 2][1] = col[1]; sp[3 * i + 2][2] = col[2]; sp[3 * i + 2][3] = 0.0f;
        keys[i] = 0xFFFFFEu - (fbits(pr.depth) >> 8u);
        visible++;
    }
    /* stable counting sort on the 24-bit key (same permutation as any stable sort) */
    uint32_t *k2 = (uint32_t *)malloc((size_t)n * 4), *v2 = (uint32_t *)malloc((size_t)n * 4);
    for (int pass = 0; pass < 3; pass++) {
        uint32_t shift = 8u * (uint32_t)pass, hist[257];
        memset(hist, 0, sizeof hist);
        for (uint32_t i = 0; i < n; i++) hist[((keys[i] >> shift) & 255u) + 1]++;
        for (int d = 0; d < 256; d++) hist[d + 1] += hist[d];
        for (uint32_t i = 0; i < n; i++) {
            uint32_t d = (keys[i] >> shift) & 255u;
            k2[hist[d]] = keys[i];
            v2[hist[d]] = vals[i];
            hist[d]++;
        }
        memcpy(keys, k2, (size_t)n * 4);
        memcpy(vals, v2, (size_t)n * 4);
    }
    /* back-to-front blend per pixel centre */
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            float cr = fr.bg[0], cg = fr.bg[1], cb = fr.bg[2];
            float qx = (float)x + 0.5f, qy = (float)y + 0.5f;
            for (uint32_t j = 0; j < visible; j++) {
                const float(*s)[4] = (const float(*)[4]) & sp[3 * vals[j]];
                float dx = qx - s[0][0], dy = qy - s[0][1];
                float u = (dx * s[1][0] + dy * s[1][1]) / (s[1][0] * s[1][0] + s[1][1] * s[1][1]);
                float v = (dx * s[1][2] + dy * s[1][3]) / (s[1][2] * s[1][2] + s[1][3] * s[1][3]);
                float r2 = u * u + v * v;
                float kk = s[0][3];
                if (r2 > kk * kk) continue;
                float alpha = om_min(s[0][2] * om_expf(-0.5f * r2), 0.99f);
                if (alpha < 0.00392156862745098f) continue;
                cr = s[2][0] * alpha + cr * (1.0f - alpha);
                cg = s[2][1] * alpha + cg * (1.0f - alpha);
                cb = s[2][2] * alpha + cb * (1.0f - alpha);
            }
            size_t i = (size_t)y * W + x;
            accum[4 * i + 0] = cr;
            accum[4 * i + 1] = cg;
            accum[4 * i + 2] = cb;
            accum[4 * i + 3] = 1.0f;
        }
    if (keys_out) memcpy(keys_out, keys, (size_t)n * 4);
    if (vals_out) memcpy(vals_out, vals, (size_t)n * 4);
    if (visible_out) *visible_out = visible;
    free(keys); free(vals); free(k2); free(v2); free(sp);
    return 0;
}

/* gps_composite.wgsl:153-179 -> RGBA8 (round to nearest even of c * 255) */
static float oetf(float c) {
    if (c <= 0.0031308f) return 12.92f * c;
    return 1.055f * om_expf((1.0f / 2.4f) * om_logf(c)) - 0.055f;
}
EXPORT void mir_composite(const float *accum, uint32_t W, uint32_t H, float exposure, uint32_t gamma_on, uint32_t ow,
                          uint32_t oh, uint8_t *out) {
    for (uint32_t Y = 0; Y < oh; Y++)
        for (uint32_t X = 0; X < ow; X++) {
            float uvx = ((float)X + 0.5f) / (float)ow, uvy = ((float)Y + 0.5f) / (float)oh;
            float qx = uvx * (float)W - 0.5f, qy = uvy * (float)H - 0.5f;
            float flx = __builtin_floorf(qx), fly = __builtin_floorf(qy);
            int px = (int)flx, py = (int)fly;
            float fx = qx - flx, fy = qy - fly;
            float lin[4][3];
            for (int t = 0; t < 4; t++) {
                int xx = px + (t & 1), yy = py + (t >> 1);
                if (xx < 0) xx = 0;
                if (xx > (int)W - 1) xx = (int)W - 1;
                if (yy < 0) yy = 0;
                if (yy > (int)H - 1) yy = (int)H - 1;
                const float *a = &accum[4 * ((size_t)yy * W + (size_t)xx)];
                float wgt = om_max(a[3], 1.0f);
                for (int k = 0; k < 3; k++) lin[t][k] = a[k] / wgt;
            }
            for (int k = 0; k < 3; k++) {
                float top = om_mix(lin[0][k], lin[1][k], fx);
                float bot = om_mix(lin[2][k], lin[3][k], fx);
                float cc = om_mix(top, bot, fy) * exposure;
                cc = om_clamp(cc, 0.0f, 1.0f);
                if (gamma_on) cc = oetf(cc);
                out[4 * ((size_t)Y * ow + X) + k] = (uint8_t)om_f2u(om_rint(cc * 255.0f));
            }
            out[4 * ((size_t)Y * ow + X) + 3] = 255u;
        }
}

/* f32 maths test hooks */
EXPORT void mir_expf_n(const float *x, float *y, size_t n) { for (size_t i = 0; i < n; i++) y[i] = om_expf(x[i]); }
EXPORT void mir_logf_n(const float *x, float *y, size_t n) { for (size_t i = 0; i < n; i++) y[i] = om_logf(x[i]); }
EXPORT void mir_sinf_n(const float *x, float *y, size_t n) { for (size_t i = 0; i < n; i++) y[i] = om_sinf(x[i]); }
EXPORT void mir_cosf_n(const float *x, float *y, size_t n) { for (size_t i = 0; i < n; i++) y[i] = om_cosf(x[i]); }

/* lib.rs:31-39 (phase-0 sample generator, p0_sample.wgsl:205-218) and the two-pass CPU reference lib.rs:169-183 */
EXPORT void mir_phase0(uint32_t seed, uint32_t pixel_base, uint32_t pixels, uint32_t samples, uint32_t *depth,
                       uint32_t *color) {
    for (uint32_t i = 0; i < pixels; i++) { depth[pixel_base + i] = 0xFFFFFFFFu; color[pixel_base + i] = 0xFFFFFFFFu; }
    for (int pass = 0; pass < 2; pass++)
        for (uint32_t i = 0; i < samples; i++) {
            uint32_t h0 = mir_pcg(i ^ mir_pcg(seed)), h1 = mir_pcg(h0), h2 = mir_pcg(h1);
            uint32_t pix = pixel_base + (h0 % pixels);
            uint32_t bits = 0x3F800000u | (h1 & 0x007FFFFFu);
            uint32_t key = (bits | 0x80000000u) & 0xFFFF0000u;
            uint32_t rgb = h2 & 0x00FFFFFFu;
            if (pass == 0) {
                if (key < depth[pix]) depth[pix] = key;
            } else if (depth[pix] == key && rgb < color[pix]) {
                color[pix] = rgb;
            }
        }
}
