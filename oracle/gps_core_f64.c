/*
 * CPU ORACLE -- test infrastructure only.  Never linked into the product
 * library; only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
 * load it.
 *
 * Plain-C f64 restatement of the reference's CPU oracle crate gps-core:
 *   crates/gps-core/src/lib.rs     (maths, RNG, calibration, radius table)
 *   crates/gps-core/src/oracle.rs  (analytic / Monte Carlo / PBVR-3D renderers,
 *                                   PSNR, SSIM)
 *   crates/gps-core/src/camera.rs  (orbit camera -> OracleCamera::look_at)
 * glam 0.29.3 (Cargo.lock) supplies the reference's f64 vector algebra; the
 * few glam routines used (quaternion -> matrix, normalise, 2x2 inverse and
 * determinant, matrix products) are restated below from glam's published
 * definitions.  Rust's f64 ln/exp/sin/cos/powf/ln_1p/exp_m1 are glibc's on
 * Linux, which this file calls directly.
 *
 * Pinned by: the six C++ fixtures of tests/fixtures (cpp_crosscheck.rs),
 * the analytic vectors of vectors.rs, and the closed-form checks of
 * tests/math.rs, restated in tests/test_oracle_f64.py.
 *
 * Build: see oracle/Makefile (-O2 -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ vectors
 * Matrices are glam-style column-major: m[c][r] = column c, row r. */
typedef struct { double x, y; } d2;
typedef struct { double x, y, z; } d3;
typedef struct { d3 c[3]; } m3;
typedef struct { d2 c[2]; } m2;

static d3 v3(double x, double y, double z) { d3 v = {x, y, z}; return v; }
static d3 add3(d3 a, d3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 sub3(d3 a, d3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 mul3(d3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross3(d3 a, d3 b) { return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
static d3 normalize3(d3 a) { return mul3(a, 1.0 / sqrt(dot3(a, a))); }
static d3 m3v(m3 m, d3 v) { return add3(add3(mul3(m.c[0], v.x), mul3(m.c[1], v.y)), mul3(m.c[2], v.z)); }
static m3 m3m(m3 a, m3 b) { m3 r; for (int i = 0; i < 3; i++) r.c[i] = m3v(a, b.c[i]); return r; }
static m3 m3t(m3 a) {
    m3 r;
    r.c[0] = v3(a.c[0].x, a.c[1].x, a.c[2].x);
    r.c[1] = v3(a.c[0].y, a.c[1].y, a.c[2].y);
    r.c[2] = v3(a.c[0].z, a.c[1].z, a.c[2].z);
    return r;
}
static double m3get(m3 m, int row, int col) { const double *c = &m.c[col].x; return c[row]; }
static d2 v2(double x, double y) { d2 v = {x, y}; return v; }
static d2 m2v(m2 m, d2 v) { return v2(m.c[0].x * v.x + m.c[1].x * v.y, m.c[0].y * v.x + m.c[1].y * v.y); }
static double m2det(m2 m) { return m.c[0].x * m.c[1].y - m.c[0].y * m.c[1].x; }
static m2 m2inv(m2 m) {
    double inv = 1.0 / m2det(m);
    m2 r;
    r.c[0] = v2(m.c[1].y * inv, -m.c[0].y * inv);
    r.c[1] = v2(-m.c[1].x * inv, m.c[0].x * inv);
    return r;
}
static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double maxd(double a, double b) { return fmax(a, b); }
static double mind(double a, double b) { return fmin(a, b); }

/* glam DMat3::from_quat of a normalised (x, y, z, w) quaternion */
static m3 quat_to_rotation(double qx, double qy, double qz, double qw) {
    double inv = 1.0 / sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
    double x = qx * inv, y = qy * inv, z = qz * inv, w = qw * inv;
    double x2 = x + x, y2 = y + y, z2 = z + z;
    double xx = x * x2, xy = x * y2, xz = x * z2;
    double yy = y * y2, yz = y * z2, zz = z * z2;
    double wx = w * x2, wy = w * y2, wz = w * z2;
    m3 r;
    r.c[0] = v3(1.0 - (yy + zz), xy + wz, xz - wy);
    r.c[1] = v3(xy - wz, 1.0 - (xx + zz), yz + wx);
    r.c[2] = v3(xz + wy, yz - wx, 1.0 - (xx + yy));
    return r;
}

/* ------------------------------------------------------------ public types */
typedef struct {
    double pos[3];
    double log_scale[3];
    double rot_xyzw[4];     /* glam order (x, y, z, w) */
    double opacity;
    double color[3];
} OrcGaussian;              /* oracle.rs:14-23 */

typedef struct {
    double view_rot[9];     /* row-major: view_rot[3*r + c] */
    double view_pos[3];
    double focal_x, focal_y, cx, cy, low_pass;
} OrcCamera;                /* oracle.rs:33-42 */

typedef struct { double points, candidates, subpixel_writes; } OrcStats; /* oracle.rs:99-104 */

typedef struct {
    uint32_t footprint_subpixels;
    double adaptive_kappa;
    uint32_t footprint_max;
    uint32_t compensate_blur;
    double density_scale;
    uint32_t per_particle_seed;
    uint32_t max_points_per_splat;
} OrcMcOptions;             /* oracle.rs:204-226 */

typedef struct {
    uint32_t level;         /* CalibrationLevel 0..3 */
    uint32_t rule;          /* 0 Proportional, 1 Extinction, 2 Dilog */
    uint32_t radial_correction;
    double base_k;
    double reference_pixel_length;
    double near_z;
    uint32_t low_pass_jitter;
    uint32_t centre_depth;
    uint32_t linearized_projection;
    double density_scale;
    uint32_t per_particle_seed;
} OrcParticleOptions;       /* oracle.rs:316-350 */

static m3 cam_rot(const OrcCamera *c) {
    m3 r;   /* column c holds view_rot[:, c] */
    for (int col = 0; col < 3; col++) r.c[col] = v3(c->view_rot[col], c->view_rot[3 + col], c->view_rot[6 + col]);
    return r;
}

/* ---------------------------------------------------------------- lib.rs */
#define SH_C0 0.28209479177387814
#define DILOG_AT_ONE 1.6449340668482264
#define MAX_OPACITY (1.0 - 1e-12)

EXPORT double orc_sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }   /* lib.rs:21-23 */

EXPORT uint32_t orc_sh_rest_per_channel(uint32_t degree) { return (degree + 1) * (degree + 1) - 1; } /* lib.rs:29-31 */

/* lib.rs:36-87 */
EXPORT void orc_eval_sh(uint32_t degree, const double dc[3], const double *rest, const double dir[3],
                        uint32_t storage_stride, double out[3]) {
    const double C1 = 0.4886025119029199;
    const double C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                          0.5462742152960396};
    const double C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                          -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
    d3 res = mul3(v3(dc[0], dc[1], dc[2]), SH_C0);
    if (degree > 3) degree = 3;
    if (degree >= 1 && rest) {
        uint32_t r = storage_stride;
        uint32_t need = orc_sh_rest_per_channel(degree);
        if (r < need) r = need;
#define SH(k) v3(rest[(k)], rest[r + (k)], rest[2 * r + (k)])
        double x = dir[0], y = dir[1], z = dir[2];
        res = add3(res, add3(add3(mul3(SH(0), -C1 * y), mul3(SH(1), C1 * z)), mul3(SH(2), -C1 * x)));
        if (degree >= 2) {
            double xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            d3 t = mul3(SH(3), C2[0] * xy);
            t = add3(t, mul3(SH(4), C2[1] * yz));
            t = add3(t, mul3(SH(5), C2[2] * (2.0 * zz - xx - yy)));
            t = add3(t, mul3(SH(6), C2[3] * xz));
            t = add3(t, mul3(SH(7), C2[4] * (xx - yy)));
            res = add3(res, t);
            if (degree >= 3) {
                d3 u = mul3(SH(8), C3[0] * y * (3.0 * xx - yy));
                u = add3(u, mul3(SH(9), C3[1] * xy * z));
                u = add3(u, mul3(SH(10), C3[2] * y * (4.0 * zz - xx - yy)));
                u = add3(u, mul3(SH(11), C3[3] * z * (2.0 * zz - 3.0 * xx - 3.0 * yy)));
                u = add3(u, mul3(SH(12), C3[4] * x * (4.0 * zz - xx - yy)));
                u = add3(u, mul3(SH(13), C3[5] * z * (xx - yy)));
                u = add3(u, mul3(SH(14), C3[6] * x * (xx - 3.0 * yy)));
                res = add3(res, u);
            }
        }
#undef SH
    }
    out[0] = maxd(res.x + 0.5, 0.0);
    out[1] = maxd(res.y + 0.5, 0.0);
    out[2] = maxd(res.z + 0.5, 0.0);
}

/* lib.rs:98-106: Sigma = R S^2 R^T */
static m3 covariance_3d(const double *ls, const double *q) {
    m3 r = quat_to_rotation(q[0], q[1], q[2], q[3]);
    m3 s2;
    memset(&s2, 0, sizeof s2);
    s2.c[0].x = exp(2.0 * ls[0]);
    s2.c[1].y = exp(2.0 * ls[1]);
    s2.c[2].z = exp(2.0 * ls[2]);
    return m3m(m3m(r, s2), m3t(r));
}

typedef struct { m3 view_rot; d3 view_pos; double focal_x, focal_y, tan_fov_x, tan_fov_y; } Pinhole; /* lib.rs:109-117 */

static d3 world_to_camera(const Pinhole *c, d3 w) { return m3v(c->view_rot, sub3(w, c->view_pos)); } /* lib.rs:132-134 */

/* lib.rs:137-165 */
static m2 covariance_2d(m3 cov3d, d3 mean_world, const Pinhole *cam, double low_pass) {
    d3 t = world_to_camera(cam, mean_world);
    double limx = 1.3 * cam->tan_fov_x, limy = 1.3 * cam->tan_fov_y;
    double txtz = t.x / t.z, tytz = t.y / t.z;
    t.x = clampd(txtz, -limx, limx) * t.z;
    t.y = clampd(tytz, -limy, limy) * t.z;
    m3 j;
    j.c[0] = v3(cam->focal_x / t.z, 0.0, 0.0);
    j.c[1] = v3(0.0, cam->focal_y / t.z, 0.0);
    j.c[2] = v3(-cam->focal_x * t.x / (t.z * t.z), -cam->focal_y * t.y / (t.z * t.z), 0.0);
    m3 tm = m3m(j, cam->view_rot);
    m3 cov = m3m(m3m(tm, cov3d), m3t(tm));
    m2 c2;
    c2.c[0] = v2(m3get(cov, 0, 0), m3get(cov, 1, 0));
    c2.c[1] = v2(m3get(cov, 0, 1), m3get(cov, 1, 1));
    c2.c[0].x += low_pass;
    c2.c[1].y += low_pass;
    return c2;
}

/* lib.rs:168-174 */
static m2 cholesky_2d(m2 a) {
    double eps = 1e-300;
    double l00 = sqrt(maxd(a.c[0].x, eps));
    double l10 = a.c[0].y / l00;
    double l11 = sqrt(maxd(a.c[1].y - l10 * l10, eps));
    m2 r;
    r.c[0] = v2(l00, l10);
    r.c[1] = v2(0.0, l11);
    return r;
}

/* lib.rs:178-190 */
static double dilog_series(double x) {
    double term = x, sum = x;
    for (int k = 2; k < 1000; k++) {
        term *= x;
        double add = term / ((double)k * (double)k);
        sum += add;
        if (fabs(add) < 1e-18 * fabs(sum)) break;
    }
    return sum;
}

/* lib.rs:193-212 */
EXPORT double orc_dilog(double x) {
    if (x >= 1.0) return DILOG_AT_ONE;
    if (x <= -1.0) return -DILOG_AT_ONE / 2.0;
    if (x == 0.0) return 0.0;
    if (x < 0.0) {
        double y = x / (x - 1.0);
        double l = log(1.0 - x);
        return -dilog_series(y) - 0.5 * l * l;
    }
    if (x <= 0.5) return dilog_series(x);
    return DILOG_AT_ONE - log(x) * log(1.0 - x) - dilog_series(1.0 - x);
}

/* lib.rs:219-248 */
EXPORT double orc_inv_dilog(double target) {
    double below_one;
    uint64_t bits;
    double d1 = DILOG_AT_ONE;
    memcpy(&bits, &d1, 8);
    bits -= 1;
    memcpy(&below_one, &bits, 8);
    target = clampd(target, 0.0, below_one);
    if (target <= 0.0) return 0.0;
    double lo = 0.0, hi = 1.0;
    double y = target / DILOG_AT_ONE;
    for (int i = 0; i < 40; i++) {
        double f = orc_dilog(y) - target;
        if (f > 0.0) hi = y; else lo = y;
        double dfdy = (y > 1e-12 && y < 1.0 - 1e-15) ? -log(1.0 - y) / y : 1.0;
        double ny = y - f / dfdy;
        if (!(ny > lo && ny < hi)) ny = 0.5 * (lo + hi);
        if (fabs(ny - y) <= 1e-14 * (1.0 + y)) return ny;
        y = ny;
    }
    return y;
}

/* lib.rs:253-258 */
EXPORT double orc_expected_point_count(double det, double opacity) {
    if (det <= 0.0 || opacity <= 0.0) return 0.0;
    return 2.0 * M_PI * sqrt(det) * orc_dilog(mind(opacity, 1.0));
}

/* lib.rs:260-280 */
EXPORT double orc_pixel_density_scale(double depth, double fx, double fy, double rpl, double near_z) {
    if (!isfinite(depth) || !isfinite(fx) || !isfinite(fy) || !isfinite(rpl) || !isfinite(near_z) ||
        depth <= maxd(near_z, 0.0) || fx <= 0.0 || fy <= 0.0 || rpl <= 0.0)
        return 0.0;
    double ratio = rpl / depth;
    return mind(ratio * ratio * fx * fy, 1.0e12);
}

/* lib.rs:287-293 */
static double projected_footprint_area(m2 c) {
    double det = c.c[0].x * c.c[1].y - c.c[1].x * c.c[0].y;
    if (!isfinite(det) || det <= 0.0) return 0.0;
    return 2.0 * M_PI * sqrt(det);
}

/* lib.rs:302-312 */
EXPORT double orc_opacity_term(uint32_t rule, double opacity) {
    if (!isfinite(opacity) || opacity <= 0.0) return 0.0;
    double o = mind(opacity, MAX_OPACITY);
    switch (rule) {
    case 0: return o;
    case 1: return -log1p(-o);
    default: return orc_dilog(o);
    }
}

typedef struct {
    double object_depth, splat_depth, focal_x, focal_y, near_z, reference_pixel_length, footprint_area, spp;
} OrcCalibInputs; /* lib.rs:322-332 */

/* lib.rs:349-391 */
EXPORT double orc_calibrated_count(uint32_t level, uint32_t rule, double opacity, double base_k, double density_scale,
                                   const OrcCalibInputs *inp) {
    double g = orc_opacity_term(rule, opacity);
    if (!(g > 0.0) || !isfinite(density_scale) || density_scale <= 0.0) return 0.0;
    double scale;
    switch (level) {
    case 0: scale = base_k; break;
    case 1: scale = base_k * orc_pixel_density_scale(inp->object_depth, inp->focal_x, inp->focal_y,
                                                     inp->reference_pixel_length, inp->near_z); break;
    case 2: scale = base_k * orc_pixel_density_scale(inp->splat_depth, inp->focal_x, inp->focal_y,
                                                     inp->reference_pixel_length, inp->near_z); break;
    default: scale = inp->spp * inp->footprint_area; break;
    }
    double n = density_scale * g * scale;
    return (isfinite(n) && n > 0.0) ? n : 0.0;
}

/* lib.rs:393-399 */
static double extinction_over_linear(double x) { return x < 1e-8 ? 1.0 + 0.5 * x : -log1p(-x) / x; }

/* lib.rs:402-410 */
EXPORT double orc_radial_keep_probability(double opacity, double r2) {
    if (!isfinite(opacity) || opacity <= 0.0) return 1.0;
    if (isnan(r2) || r2 < 0.0) r2 = 0.0;
    double o = mind(opacity, MAX_OPACITY);
    double y = o * exp(-0.5 * r2);
    return clampd(extinction_over_linear(y) / extinction_over_linear(o), 0.0, 1.0);
}

/* lib.rs:412-418 */
static double footprint_point_count(double expected, uint32_t s) {
    if (!isfinite(expected) || expected <= 0.0) return 0.0;
    double sd = (double)(s > 1 ? s : 1);
    return expected / (sd * sd);
}

/* lib.rs:420-433 */
static uint32_t adaptive_footprint(double sigma, double kappa, uint32_t s_max) {
    if (!(kappa > 0.0) || !isfinite(sigma) || sigma <= 0.0) return 1;
    uint32_t hi = s_max > 1 ? s_max : 1;
    double s = floor(kappa * sigma);
    if (!(s >= 1.0)) return 1;
    if (s >= (double)hi) return hi;
    return (uint32_t)s;
}

/* lib.rs:438-442 */
EXPORT uint32_t orc_pcg_hash(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28) + 4)) ^ s) * 277803737u;
    return (w >> 22) ^ w;
}

/* lib.rs:450-465; mode 0 = Deterministic, 1 = FrameVarying */
EXPORT uint32_t orc_particle_seed(uint32_t mode, uint32_t splat_id, uint32_t k, uint32_t frame) {
    uint32_t s = orc_pcg_hash(splat_id * 747796405u + k + 1u);
    if (mode == 1) s = orc_pcg_hash(s ^ (frame * 2654435761u + 1u));
    return s ? s : 1u;
}

/* lib.rs:469-498 */
typedef struct { uint32_t state; int has_spare; double spare; } RandStream;
static RandStream rs_new(uint32_t seed) { RandStream r = {seed ? seed : 0x9E3779B9u, 0, 0.0}; return r; }
static double rs_next(RandStream *r) { r->state = orc_pcg_hash(r->state); return (double)r->state * (1.0 / 4294967296.0); }
static double rs_normal(RandStream *r) {
    if (r->has_spare) { r->has_spare = 0; return r->spare; }
    double u1 = maxd(rs_next(r), 1e-300);
    double u2 = rs_next(r);
    double rad = sqrt(-2.0 * log(u1));
    double th = 2.0 * M_PI * u2;
    r->spare = rad * sin(th);
    r->has_spare = 1;
    return rad * cos(th);
}

/* lib.rs:500-519 */
static int32_t poisson_sample(double lambda, RandStream *rng) {
    if (lambda <= 0.0) return 0;
    if (lambda < 30.0) {
        double l = exp(-lambda);
        int32_t k = 0;
        double p = 1.0;
        for (;;) {
            k += 1;
            p *= rs_next(rng);
            if (p <= l) break;
        }
        return k - 1;
    }
    double n = round(lambda + sqrt(lambda) * rs_normal(rng));
    n = maxd(n, 0.0);
    return n >= 2147483647.0 ? 2147483647 : (int32_t)n;
}

EXPORT double orc_poisson_mean(double lambda, uint32_t seed, uint32_t n) {
    RandStream r = rs_new(seed);
    double s = 0.0;
    for (uint32_t i = 0; i < n; i++) s += poisson_sample(lambda, &r);
    return s / n;
}

/* lib.rs:522-536 */
EXPORT double orc_sample_corrected_radius(double opacity, double u) {
    double o = clampd(opacity, 0.0, 1.0);
    if (o <= 0.0) return 0.0;
    u = clampd(u, 0.0, 1.0);
    double d = orc_dilog(o);
    double target = (1.0 - u) * d;
    double y = mind(orc_inv_dilog(target), o);
    if (y <= 0.0) return 40.0;
    double r2 = -2.0 * log(y / o);
    return sqrt(maxd(r2, 0.0));
}

/* lib.rs:538-542 */
static d2 sample_corrected_offset(double opacity, m2 cov2d, RandStream *rng) {
    double r = orc_sample_corrected_radius(opacity, rs_next(rng));
    double theta = 2.0 * M_PI * rs_next(rng);
    return m2v(cholesky_2d(cov2d), v2(r * cos(theta), r * sin(theta)));
}

typedef struct { d2 mean; m2 cov; double opacity; d3 color; double depth; } G2; /* lib.rs:546-553 */

/* lib.rs:555-559 */
static double alpha_at(const G2 *g, d2 p) {
    d2 d = v2(p.x - g->mean.x, p.y - g->mean.y);
    d2 t = m2v(m2inv(g->cov), d);
    double m = d.x * t.x + d.y * t.y;
    return clampd(g->opacity, 0.0, 1.0) * exp(-0.5 * m);
}

/* lib.rs:561-564 */
EXPORT double orc_expected_coverage(double alpha, double pixel_area) {
    alpha = clampd(alpha, 0.0, 1.0 - 1e-15);
    return 1.0 - pow(1.0 - alpha, pixel_area);
}

static int cmp_depth(const void *a, const void *b) {
    const G2 *x = (const G2 *)a, *y = (const G2 *)b;
    if (x->depth < y->depth) return -1;
    if (x->depth > y->depth) return 1;
    return 0;
}

/* stable insertion sort by depth (Rust sort_by is stable) */
static void sort_layers(G2 *l, size_t n) {
    for (size_t i = 1; i < n; i++) {
        G2 t = l[i];
        size_t j = i;
        while (j > 0 && cmp_depth(&l[j - 1], &t) > 0) { l[j] = l[j - 1]; j--; }
        l[j] = t;
    }
}

/* lib.rs:567-585; `layers` must already be depth-sorted */
static d3 composite_expected_sorted(const G2 *layers, size_t n, d2 pixel, d3 bg, double pixel_area) {
    d3 acc = v3(0, 0, 0);
    double t = 1.0;
    for (size_t i = 0; i < n; i++) {
        double a = orc_expected_coverage(alpha_at(&layers[i], pixel), pixel_area);
        acc = add3(acc, mul3(layers[i].color, t * a));
        t *= 1.0 - a;
        if (t <= 1e-9) break;
    }
    return add3(acc, mul3(bg, t));
}

/* lib.rs:590-599 */
EXPORT uint32_t orc_ordered_depth_key(float depth) {
    uint32_t bits;
    memcpy(&bits, &depth, 4);
    uint32_t mask = (bits & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;
    return bits ^ mask;
}
EXPORT float orc_ordered_depth_key_inverse(uint32_t key) {
    uint32_t mask = (key & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu;
    uint32_t b = key ^ mask;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

/* lib.rs:615-636 */
#define LUT_O 65
#define LUT_V 129
#define LUT_VMAX 17.0
EXPORT void orc_radius_lut(float *out) {
    for (int i = 0; i < LUT_O; i++) {
        double x = (double)i / (double)(LUT_O - 1);
        double o = clampd(1.0 - (1.0 - x) * (1.0 - x), 1e-6, 1.0);
        for (int j = 0; j < LUT_V; j++) {
            double w = (double)j * sqrt(LUT_VMAX) / (double)(LUT_V - 1);
            double u = -expm1(-w * w);
            out[i * LUT_V + j] = (float)orc_sample_corrected_radius(o, u);
        }
    }
}

/* lib.rs:639-648 */
EXPORT double orc_radius_from_lut(const float *lut, double o, double u) {
    double v = -log(maxd(1.0 - u, 6e-8));
    double fo = (1.0 - sqrt(1.0 - clampd(o, 0.0, 1.0))) * (double)(LUT_O - 1);
    double fv = clampd(sqrt(v) / sqrt(LUT_VMAX), 0.0, 1.0) * (double)(LUT_V - 1);
    size_t i0 = (size_t)fo, j0 = (size_t)fv;
    if (i0 > LUT_O - 2) i0 = LUT_O - 2;
    if (j0 > LUT_V - 2) j0 = LUT_V - 2;
    double a = fo - (double)i0, b = fv - (double)j0;
#define AT(i, j) ((double)lut[(i) * LUT_V + (j)])
    double r = (1.0 - a) * ((1.0 - b) * AT(i0, j0) + b * AT(i0, j0 + 1)) +
               a * ((1.0 - b) * AT(i0 + 1, j0) + b * AT(i0 + 1, j0 + 1));
#undef AT
    return maxd(r, 0.0);
}

/* ------------------------------------------------------------- oracle.rs */
#define OFFSCREEN_SIGMA 5.9

/* oracle.rs:53-70 */
EXPORT void orc_look_at(const double eye[3], const double target[3], const double up[3], double fov_y, uint32_t w,
                        uint32_t h, OrcCamera *out) {
    d3 e = v3(eye[0], eye[1], eye[2]);
    d3 f = normalize3(sub3(v3(target[0], target[1], target[2]), e));
    d3 right = normalize3(cross3(f, v3(up[0], up[1], up[2])));
    d3 down = cross3(f, right);
    /* rows of view_rot are the camera axes */
    const d3 rows[3] = {right, down, f};
    for (int r = 0; r < 3; r++) {
        out->view_rot[3 * r + 0] = rows[r].x;
        out->view_rot[3 * r + 1] = rows[r].y;
        out->view_rot[3 * r + 2] = rows[r].z;
    }
    out->view_pos[0] = eye[0];
    out->view_pos[1] = eye[1];
    out->view_pos[2] = eye[2];
    double focal = 0.5 * (double)h / tan(0.5 * fov_y);
    out->focal_x = focal;
    out->focal_y = focal;
    out->cx = 0.5 * (double)w;
    out->cy = 0.5 * (double)h;
    out->low_pass = 0.3;
}

/* camera.rs:31-40 (OrbitCamera::eye) + camera.rs:286-288 (to_camera) */
EXPORT void orc_orbit_camera(const double target[3], double distance, double yaw, double pitch, const double up_in[3],
                             double fov_y, uint32_t w, uint32_t h, OrcCamera *out) {
    d3 up = normalize3(v3(up_in[0], up_in[1], up_in[2]));
    d3 helper = fabs(up.x) < 0.9 ? v3(1, 0, 0) : v3(0, 0, 1);
    d3 a = normalize3(cross3(up, helper));
    d3 b = cross3(up, a);
    d3 horiz = add3(mul3(a, cos(yaw)), mul3(b, sin(yaw)));
    d3 dir = add3(mul3(horiz, cos(pitch)), mul3(up, sin(pitch)));
    d3 eye = add3(v3(target[0], target[1], target[2]), mul3(dir, distance));
    double e[3] = {eye.x, eye.y, eye.z};
    orc_look_at(e, target, up_in, fov_y, w, h, out);
}

static Pinhole to_pinhole(const OrcCamera *c) {
    Pinhole p;
    p.view_rot = cam_rot(c);
    p.view_pos = v3(c->view_pos[0], c->view_pos[1], c->view_pos[2]);
    p.focal_x = c->focal_x;
    p.focal_y = c->focal_y;
    p.tan_fov_x = 4.0;
    p.tan_fov_y = 4.0;
    return p;
}

static OrcCamera with_default_pp(OrcCamera c, size_t w, size_t h) {
    if (c.cx == 0.0 && c.cy == 0.0) {
        c.cx = 0.5 * (double)w;
        c.cy = 0.5 * (double)h;
    }
    return c;
}

typedef struct { d2 mean; m2 cov2d; double det, depth, opacity; d3 color; } Projected; /* oracle.rs:106-114 */

/* oracle.rs:116-136 */
static int project(const OrcGaussian *g, const OrcCamera *c, Projected *out) {
    Pinhole pin = to_pinhole(c);
    d3 pos = v3(g->pos[0], g->pos[1], g->pos[2]);
    d3 cm = world_to_camera(&pin, pos);
    if (cm.z <= 1e-4) return 0;
    m3 cov3 = covariance_3d(g->log_scale, g->rot_xyzw);
    m2 cov2 = covariance_2d(cov3, pos, &pin, c->low_pass);
    double det = m2det(cov2);
    if (det <= 0.0) return 0;
    out->mean = v2(c->focal_x * cm.x / cm.z + c->cx, c->focal_y * cm.y / cm.z + c->cy);
    out->cov2d = cov2;
    out->det = det;
    out->depth = cm.z;
    out->opacity = clampd(g->opacity, 0.0, 1.0);
    out->color = v3(g->color[0], g->color[1], g->color[2]);
    return 1;
}

/* oracle.rs:147-153 */
static void depth_test(size_t idx, double d, d3 c, double *depth, d3 *color, OrcStats *st) {
    st->subpixel_writes += 1.0;
    if (d < depth[idx]) {
        depth[idx] = d;
        color[idx] = c;
    }
}

/* oracle.rs:156-173 */
static void write_block(d2 pixel, int64_t s, size_t w, size_t h, int64_t spp_side, double d, d3 c, double *depth,
                        d3 *color, OrcStats *st) {
    int64_t spp = spp_side * spp_side;
    int64_t gx0 = (int64_t)floor(pixel.x * (double)spp_side - 0.5 * (double)s + 0.5);
    int64_t gy0 = (int64_t)floor(pixel.y * (double)spp_side - 0.5 * (double)s + 0.5);
    int64_t gw = (int64_t)w * spp_side, gh = (int64_t)h * spp_side;
    int64_t y1 = gh < gy0 + s ? gh : gy0 + s;
    int64_t x1 = gw < gx0 + s ? gw : gx0 + s;
    for (int64_t gy = gy0 > 0 ? gy0 : 0; gy < y1; gy++) {
        for (int64_t gx = gx0 > 0 ? gx0 : 0; gx < x1; gx++) {
            int64_t px = gx / spp_side, py = gy / spp_side, sx = gx % spp_side, sy = gy % spp_side;
            size_t idx = (size_t)(((int64_t)(py * (int64_t)w + px)) * spp + sy * spp_side + sx);
            depth_test(idx, d, c, depth, color, st);
        }
    }
}

/* oracle.rs:175-186 */
static void resolve_set(size_t w, size_t h, size_t spp, const double *depth, const d3 *color, d3 bg, d3 *accum) {
    for (size_t y = 0; y < h; y++) {
        for (size_t x = 0; x < w; x++) {
            size_t base = (y * w + x) * spp;
            d3 sum = v3(0, 0, 0);
            for (size_t k = 0; k < spp; k++) sum = add3(sum, depth[base + k] < 1e299 ? color[base + k] : bg);
            accum[y * w + x] = add3(accum[y * w + x], v3(sum.x / (double)spp, sum.y / (double)spp, sum.z / (double)spp));
        }
    }
}

/* oracle.rs:190-200 */
EXPORT void orc_render_analytic(const OrcGaussian *scene, size_t n, const OrcCamera *cam_in, size_t w, size_t h,
                                const double bg[3], double *out /* w*h*3 */) {
    OrcCamera cam = with_default_pp(*cam_in, w, h);
    G2 *layers = (G2 *)malloc(sizeof(G2) * (n ? n : 1));
    size_t m = 0;
    for (size_t i = 0; i < n; i++) {
        Projected p;
        if (!project(&scene[i], &cam, &p)) continue;
        layers[m].mean = p.mean;
        layers[m].cov = p.cov2d;
        layers[m].opacity = p.opacity;
        layers[m].color = p.color;
        layers[m].depth = p.depth;
        m++;
    }
    sort_layers(layers, m);
    d3 b = v3(bg[0], bg[1], bg[2]);
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            d3 c = composite_expected_sorted(layers, m, v2((double)x + 0.5, (double)y + 0.5), b, 1.0);
            out[3 * (y * w + x) + 0] = c.x;
            out[3 * (y * w + x) + 1] = c.y;
            out[3 * (y * w + x) + 2] = c.z;
        }
    free(layers);
}

/* Supersampled analytic reference of statistics.rs:134-160 (N x N samples per pixel). */
EXPORT void orc_render_analytic_area(const OrcGaussian *scene, size_t n, const OrcCamera *cam_in, size_t w, size_t h,
                                     const double bg[3], uint32_t ss, double *out) {
    OrcCamera cam = with_default_pp(*cam_in, w, h);
    G2 *layers = (G2 *)malloc(sizeof(G2) * (n ? n : 1));
    size_t m = 0;
    for (size_t i = 0; i < n; i++) {
        Projected p;
        if (!project(&scene[i], &cam, &p)) continue;
        layers[m].mean = p.mean;
        layers[m].cov = p.cov2d;
        layers[m].opacity = p.opacity;
        layers[m].color = p.color;
        layers[m].depth = p.depth;
        m++;
    }
    sort_layers(layers, m);
    d3 b = v3(bg[0], bg[1], bg[2]);
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            d3 acc = v3(0, 0, 0);
            for (uint32_t j = 0; j < ss; j++)
                for (uint32_t i = 0; i < ss; i++) {
                    d2 p = v2((double)x + ((double)i + 0.5) / (double)ss, (double)y + ((double)j + 0.5) / (double)ss);
                    acc = add3(acc, composite_expected_sorted(layers, m, p, b, 1.0));
                }
            double k = (double)(ss * ss);
            out[3 * (y * w + x) + 0] = acc.x / k;
            out[3 * (y * w + x) + 1] = acc.y / k;
            out[3 * (y * w + x) + 2] = acc.z / k;
        }
    free(layers);
}

/* oracle.rs:229-312 */
EXPORT void orc_render_monte_carlo(const OrcGaussian *scene, size_t n, const OrcCamera *cam_in, size_t w, size_t h,
                                   size_t spp_side, size_t num_sets, const double bgv[3], uint32_t base_seed,
                                   const OrcMcOptions *opt, double *out, OrcStats *stats) {
    OrcCamera cam = with_default_pp(*cam_in, w, h);
    if (spp_side < 1) spp_side = 1;
    if (spp_side > 4) spp_side = 4;
    size_t spp = spp_side * spp_side, n_sub = w * h * spp;
    OrcStats st = {0, 0, 0};
    Projected *proj = (Projected *)malloc(sizeof(Projected) * (n ? n : 1));
    size_t m = 0;
    for (size_t i = 0; i < n; i++)
        if (project(&scene[i], &cam, &proj[m])) m++;
    uint32_t *fpv = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    m2 *scov = (m2 *)malloc(sizeof(m2) * (m ? m : 1));
    for (size_t g = 0; g < m; g++) {
        Projected *pr = &proj[g];
        scov[g] = pr->cov2d;
        uint32_t fp = opt->footprint_subpixels > 1 ? opt->footprint_subpixels : 1;
        if (opt->adaptive_kappa > 0.0) {
            double tr = pr->cov2d.c[0].x + pr->cov2d.c[1].y;
            double disc = sqrt(maxd(0.25 * tr * tr - pr->det, 0.0));
            double smin = sqrt(maxd(0.5 * tr - disc, 0.0));
            fp = adaptive_footprint(smin * (double)spp_side, opt->adaptive_kappa, opt->footprint_max);
        }
        fpv[g] = fp;
        if (opt->compensate_blur && fp > 1) {
            double v = ((double)fp / (double)spp_side) * ((double)fp / (double)spp_side) / 12.0;
            m2 sh = pr->cov2d;
            sh.c[0].x -= v;
            sh.c[1].y -= v;
            double t = sh.c[0].x + sh.c[1].y;
            double d = m2det(sh);
            if (d > 0.0 && t > 0.0 && 0.5 * t - sqrt(maxd(0.25 * t * t - d, 0.0)) >= v) scov[g] = sh;
        }
    }
    d3 *accum = (d3 *)calloc(w * h, sizeof(d3));
    double *depth = (double *)malloc(sizeof(double) * n_sub);
    d3 *color = (d3 *)malloc(sizeof(d3) * n_sub);
    d3 bg = v3(bgv[0], bgv[1], bgv[2]);
    for (size_t s = 0; s < num_sets; s++) {
        for (size_t i = 0; i < n_sub; i++) { depth[i] = 1e300; color[i] = v3(0, 0, 0); }
        uint32_t set_seed = base_seed + (uint32_t)s;
        for (size_t gid = 0; gid < m; gid++) {
            Projected *pr = &proj[gid];
            uint32_t fp = fpv[gid];
            double en = footprint_point_count(orc_expected_point_count(pr->det, pr->opacity) * (double)spp *
                                                  opt->density_scale, fp);
            if (en <= 0.0) continue;
            if (fp == 1) {
                double ex = OFFSCREEN_SIGMA * sqrt(scov[gid].c[0].x), ey = OFFSCREEN_SIGMA * sqrt(scov[gid].c[1].y);
                if (pr->mean.x + ex < 0.0 || pr->mean.x - ex >= (double)w || pr->mean.y + ey < 0.0 ||
                    pr->mean.y - ey >= (double)h)
                    continue;
            }
            RandStream rng = rs_new(orc_particle_seed(1, (uint32_t)gid, 0, set_seed));
            int32_t n_full = poisson_sample(en, &rng);
            st.points += (double)n_full;
            uint32_t nn = (uint32_t)n_full;
            if (nn > opt->max_points_per_splat) nn = opt->max_points_per_splat;
            for (uint32_t j = 0; j < nn; j++) {
                if (opt->per_particle_seed) rng = rs_new(orc_particle_seed(1, (uint32_t)gid, j + 1, set_seed));
                d2 off = sample_corrected_offset(pr->opacity, scov[gid], &rng);
                d2 pixel = v2(pr->mean.x + off.x, pr->mean.y + off.y);
                if (fp == 1) {
                    if (pixel.x < 0.0 || pixel.x >= (double)w || pixel.y < 0.0 || pixel.y >= (double)h) continue;
                    size_t px = (size_t)pixel.x, py = (size_t)pixel.y;
                    size_t sx = (size_t)((pixel.x - (double)px) * (double)spp_side);
                    size_t sy = (size_t)((pixel.y - (double)py) * (double)spp_side);
                    if (sx > spp_side - 1) sx = spp_side - 1;
                    if (sy > spp_side - 1) sy = spp_side - 1;
                    depth_test((py * w + px) * spp + sy * spp_side + sx, pr->depth, pr->color, depth, color, &st);
                } else {
                    write_block(pixel, (int64_t)fp, w, h, (int64_t)spp_side, pr->depth, pr->color, depth, color, &st);
                }
            }
        }
        resolve_set(w, h, spp, depth, color, bg, accum);
    }
    for (size_t i = 0; i < w * h; i++) {
        out[3 * i + 0] = accum[i].x / (double)num_sets;
        out[3 * i + 1] = accum[i].y / (double)num_sets;
        out[3 * i + 2] = accum[i].z / (double)num_sets;
    }
    if (stats) *stats = st;
    free(proj); free(fpv); free(scov); free(accum); free(depth); free(color);
}

/* oracle.rs:352-361 */
static m3 cholesky_3d(m3 c) {
    double a00 = m3get(c, 0, 0), a10 = m3get(c, 1, 0), a20 = m3get(c, 2, 0);
    double a11 = m3get(c, 1, 1), a21 = m3get(c, 2, 1), a22 = m3get(c, 2, 2);
    double l00 = sqrt(maxd(a00, 1e-30));
    double l10 = a10 / l00, l20 = a20 / l00;
    double l11 = sqrt(maxd(a11 - l10 * l10, 1e-30));
    double l21 = (a21 - l20 * l10) / l11;
    double l22 = sqrt(maxd(a22 - l20 * l20 - l21 * l21, 1e-30));
    m3 r;
    r.c[0] = v3(l00, l10, l20);
    r.c[1] = v3(0.0, l11, l21);
    r.c[2] = v3(0.0, 0.0, l22);
    return r;
}

typedef struct { uint32_t index; d3 pos; Projected pr; m3 l3; m2 cinv; m3 jw; double lambda; } Prep; /* oracle.rs:363-371 */

/* oracle.rs:374-499 */
EXPORT void orc_render_particles_3d(const OrcGaussian *scene, size_t n, const OrcCamera *cam_in, size_t w, size_t h,
                                    size_t spp_side, size_t num_sets, const double bgv[3], uint32_t base_seed,
                                    const OrcParticleOptions *opt, double *out, OrcStats *stats) {
    OrcCamera cam = with_default_pp(*cam_in, w, h);
    if (spp_side < 1) spp_side = 1;
    if (spp_side > 4) spp_side = 4;
    size_t spp = spp_side * spp_side, n_sub = w * h * spp;
    Pinhole pin = to_pinhole(&cam);
    OrcStats st = {0, 0, 0};
    d3 centre = v3(0, 0, 0);
    for (size_t i = 0; i < n; i++) centre = add3(centre, v3(scene[i].pos[0], scene[i].pos[1], scene[i].pos[2]));
    if (n) centre = v3(centre.x / (double)n, centre.y / (double)n, centre.z / (double)n);
    double object_depth = world_to_camera(&pin, centre).z;
    Prep *preps = (Prep *)malloc(sizeof(Prep) * (n ? n : 1));
    size_t m = 0;
    m3 wt = m3t(cam_rot(&cam));
    for (size_t si = 0; si < n; si++) {
        Projected pr;
        if (!project(&scene[si], &cam, &pr)) continue;
        const OrcGaussian *g = &scene[si];
        d3 pos = v3(g->pos[0], g->pos[1], g->pos[2]);
        m3 cov3 = covariance_3d(g->log_scale, g->rot_xyzw);
        m3 l3 = cholesky_3d(cov3);
        d3 cm = world_to_camera(&pin, pos);
        double iz = 1.0 / cm.z;
        d3 jx = v3(cam.focal_x * iz, 0.0, -cam.focal_x * cm.x * iz * iz);
        d3 jy = v3(0.0, cam.focal_y * iz, -cam.focal_y * cm.y * iz * iz);
        d3 rx = m3v(wt, jx), ry = m3v(wt, jy);
        m3 jw;
        jw.c[0] = v3(rx.x, ry.x, 0.0);
        jw.c[1] = v3(rx.y, ry.y, 0.0);
        jw.c[2] = v3(rx.z, ry.z, 0.0);
        OrcCalibInputs inp = {object_depth, cm.z, cam.focal_x, cam.focal_y, opt->near_z, opt->reference_pixel_length,
                              projected_footprint_area(pr.cov2d), (double)spp};
        double lambda = orc_calibrated_count(opt->level, opt->rule, pr.opacity, opt->base_k, opt->density_scale, &inp);
        Prep *p = &preps[m++];
        p->index = (uint32_t)si;
        p->pos = pos;
        p->pr = pr;
        p->l3 = l3;
        p->cinv = m2inv(pr.cov2d);
        p->jw = jw;
        p->lambda = lambda;
    }
    d3 *accum = (d3 *)calloc(w * h, sizeof(d3));
    double *depth = (double *)malloc(sizeof(double) * n_sub);
    d3 *color = (d3 *)malloc(sizeof(d3) * n_sub);
    d3 forward = m3t(cam_rot(&cam)).c[2];
    double jitter = opt->low_pass_jitter ? sqrt(maxd(cam.low_pass, 0.0)) : 0.0;
    d3 bg = v3(bgv[0], bgv[1], bgv[2]);
    for (size_t s = 0; s < num_sets; s++) {
        for (size_t i = 0; i < n_sub; i++) { depth[i] = 1e300; color[i] = v3(0, 0, 0); }
        for (size_t gid = 0; gid < m; gid++) {
            Prep *p = &preps[gid];
            if (p->lambda <= 0.0) continue;
            uint32_t set_seed = base_seed + (uint32_t)s;
            RandStream rng = opt->per_particle_seed ? rs_new(orc_particle_seed(1, p->index, 0, set_seed))
                                                    : rs_new(orc_particle_seed(1, (uint32_t)gid, 1, set_seed));
            int32_t nn = poisson_sample(p->lambda, &rng);
            st.candidates += (double)nn;
            for (int32_t k = 0; k < nn; k++) {
                if (opt->per_particle_seed) rng = rs_new(orc_particle_seed(1, p->index, (uint32_t)k + 1, set_seed));
                double uz = rs_normal(&rng);
                double uy = rs_normal(&rng);
                double ux = rs_normal(&rng);
                d3 u = v3(ux, uy, uz);
                d3 world = add3(p->pos, m3v(p->l3, u));
                d2 pixel;
                double d;
                if (opt->linearized_projection) {
                    d3 off = m3v(p->jw, sub3(world, p->pos));
                    pixel = v2(p->pr.mean.x + off.x, p->pr.mean.y + off.y);
                    d = p->pr.depth + dot3(forward, sub3(world, p->pos));
                } else {
                    d3 cp = world_to_camera(&pin, world);
                    if (cp.z <= opt->near_z) continue;
                    pixel = v2(cam.focal_x * cp.x / cp.z + cam.cx, cam.focal_y * cp.y / cp.z + cam.cy);
                    d = cp.z;
                }
                if (jitter > 0.0) {
                    double jy = rs_normal(&rng);
                    double jx = rs_normal(&rng);
                    pixel = v2(pixel.x + jitter * jx, pixel.y + jitter * jy);
                }
                if (opt->radial_correction) {
                    d2 dv = v2(pixel.x - p->pr.mean.x, pixel.y - p->pr.mean.y);
                    d2 t = m2v(p->cinv, dv);
                    double r2 = dv.x * t.x + dv.y * t.y;
                    if (rs_next(&rng) >= orc_radial_keep_probability(p->pr.opacity, r2)) continue;
                }
                if (opt->centre_depth) d = p->pr.depth;
                st.points += 1.0;
                if (pixel.x < 0.0 || pixel.x >= (double)w || pixel.y < 0.0 || pixel.y >= (double)h) continue;
                size_t px = (size_t)pixel.x, py = (size_t)pixel.y;
                size_t sx = (size_t)((pixel.x - (double)px) * (double)spp_side);
                size_t sy = (size_t)((pixel.y - (double)py) * (double)spp_side);
                if (sx > spp_side - 1) sx = spp_side - 1;
                if (sy > spp_side - 1) sy = spp_side - 1;
                depth_test((py * w + px) * spp + sy * spp_side + sx, d, p->pr.color, depth, color, &st);
            }
        }
        resolve_set(w, h, spp, depth, color, bg, accum);
    }
    size_t ns = num_sets ? num_sets : 1;
    for (size_t i = 0; i < w * h; i++) {
        out[3 * i + 0] = accum[i].x / (double)ns;
        out[3 * i + 1] = accum[i].y / (double)ns;
        out[3 * i + 2] = accum[i].z / (double)ns;
    }
    if (stats) *stats = st;
    free(preps); free(accum); free(depth); free(color);
}

/* oracle.rs:503-514; images are w*h*3 doubles */
EXPORT double orc_psnr(const double *a, const double *b, size_t n_pixels) {
    if (n_pixels == 0) return 0.0;
    double s = 0.0;
    for (size_t i = 0; i < n_pixels; i++) {
        double dx = a[3 * i] - b[3 * i], dy = a[3 * i + 1] - b[3 * i + 1], dz = a[3 * i + 2] - b[3 * i + 2];
        s += dx * dx + dy * dy + dz * dz;
    }
    double mse = s / (double)(n_pixels * 3);
    return mse <= 1e-20 ? 120.0 : 10.0 * log10(1.0 / mse);
}

/* oracle.rs:517-559 */
EXPORT double orc_ssim(const double *a, const double *b, size_t w, size_t h) {
#define LUM(img, i) (0.299 * (img)[3 * (i)] + 0.587 * (img)[3 * (i) + 1] + 0.114 * (img)[3 * (i) + 2])
    const double c1 = 0.01 * 0.01, c2 = 0.03 * 0.03;
    double total = 0.0;
    size_t count = 0;
    for (size_t y0 = 0; y0 + 8 <= h; y0 += 4)
        for (size_t x0 = 0; x0 + 8 <= w; x0 += 4) {
            double ma = 0, mb = 0;
            for (size_t y = 0; y < 8; y++)
                for (size_t x = 0; x < 8; x++) {
                    size_t i = (y0 + y) * w + x0 + x;
                    ma += LUM(a, i);
                    mb += LUM(b, i);
                }
            ma /= 64.0;
            mb /= 64.0;
            double va = 0, vb = 0, cov = 0;
            for (size_t y = 0; y < 8; y++)
                for (size_t x = 0; x < 8; x++) {
                    size_t i = (y0 + y) * w + x0 + x;
                    double da = LUM(a, i) - ma, db = LUM(b, i) - mb;
                    va += da * da;
                    vb += db * db;
                    cov += da * db;
                }
            va /= 63.0;
            vb /= 63.0;
            cov /= 63.0;
            total += ((2.0 * ma * mb + c1) * (2.0 * cov + c2)) / ((ma * ma + mb * mb + c1) * (va + vb + c2));
            count++;
        }
#undef LUM
    return count ? total / (double)count : 1.0;
}
