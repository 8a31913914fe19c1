/* Host sanitizer driver for the oracle (test infrastructure): built with -fsanitize=address,undefined together
 * with oracle/gps_core_f64.c and oracle/gps_mirror.c by tests/test_sanitizers.py, it runs the oracle's renderers
 * on the inputs that test writes and writes their outputs, which the test compares with the uninstrumented
 * liboracle.so.  Struct arguments are opaque byte blocks laid out by the test (oracle/oracle.py's ctypes types).
 *
 * input:  u32 n, sh_degree, w, h, first_seed, count, param_sets, params_size, orc_size, popt_size, mopt_size;
 *         n x 64-byte splats; n x 3 x ((sh_degree + 1)^2 - 1) f32 SH rest; 136-byte camera;
 *         param_sets x params_size render params; n x orc_size f64 splats; popt_size and mopt_size option blocks
 * output: per parameter set: accum, 16 u64 stats, counts, offsets, depth|id, total, skipped; then the sorted
 *         alpha path (accum, stats, keys, values, visible), the 2x composite, and the f64 renderers' images and
 *         counters (particles 3D, Monte Carlo, analytic) */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int mir_render_ensembles(const void *g, uint32_t n, const float *sh, uint32_t deg, const float *lut, const void *cam,
                         const void *p, uint32_t W, uint32_t H, uint32_t spp, uint32_t first_seed, uint32_t count,
                         float *accum, uint64_t *stats, uint32_t *counts, uint32_t *offsets, uint64_t *depth_id,
                         uint32_t *total, uint32_t *skipped);
int mir_render_sorted(const void *g, uint32_t n, const float *sh, uint32_t deg, const void *cam, const void *p,
                      uint32_t W, uint32_t H, uint32_t spp, float *accum, uint64_t *stats, uint32_t *keys,
                      uint32_t *vals, uint32_t *vis);
void mir_composite(const float *accum, uint32_t W, uint32_t H, float exposure, uint32_t gamma_on, uint32_t ow,
                   uint32_t oh, uint8_t *out);
void mir_set_threads(int n);
void orc_radius_lut(float *out);
void orc_render_particles_3d(const void *scene, size_t n, const void *cam, size_t w, size_t h, size_t spp,
                             size_t sets, const double *bg, uint32_t seed, const void *opt, double *out, double *st);
void orc_render_monte_carlo(const void *scene, size_t n, const void *cam, size_t w, size_t h, size_t spp,
                            size_t sets, const double *bg, uint32_t seed, const void *opt, double *out, double *st);
void orc_render_analytic(const void *scene, size_t n, const void *cam, size_t w, size_t h, const double *bg,
                         double *out);

static void *rd(FILE *f, size_t bytes) {
    void *p = calloc(bytes ? bytes : 1, 1);
    if (!p || fread(p, 1, bytes, f) != bytes) { fprintf(stderr, "short input\n"); exit(2); }
    return p;
}
static void wr(FILE *f, const void *p, size_t bytes) {
    if (fwrite(p, 1, bytes, f) != bytes) { fprintf(stderr, "short write\n"); exit(2); }
}

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
    if (!in || !out) return 2;
    uint32_t *hd = rd(in, 11 * 4);
    const uint32_t n = hd[0], deg = hd[1], W = hd[2], H = hd[3], seed = hd[4], count = hd[5], sets = hd[6];
    const size_t psz = hd[7], osz = hd[8], popt = hd[9], mopt = hd[10];
    const uint32_t rest = (deg + 1) * (deg + 1) - 1;
    void *splats = rd(in, (size_t)n * 64);
    float *sh = rd(in, (size_t)n * 3 * rest * 4);
    void *cam = rd(in, 136);
    unsigned char *params = rd(in, sets * psz);
    void *orc = rd(in, (size_t)n * osz);
    void *po = rd(in, popt), *mo = rd(in, mopt);
    fclose(in);
    mir_set_threads(2);
    float *lut = calloc(65 * 129, 4);
    orc_radius_lut(lut);
    const size_t X = (size_t)W * H;
    for (uint32_t s = 0; s < sets; s++) {
        const void *p = params + s * psz;
        const uint32_t spp = *(const uint32_t *)p;
        const size_t S = X * spp * spp;
        float *accum = calloc(X * 4, 4);
        uint64_t stats[16] = {0}, *depth = calloc(S, 8);
        uint32_t *counts = calloc(n ? n : 1, 4), *offsets = calloc(n ? n : 1, 4), total = 0, skipped = 0;
        int rc = mir_render_ensembles(splats, n, rest ? sh : NULL, rest ? deg : 0, lut, cam, p, W, H, spp, seed,
                                      count, accum, stats, counts, offsets, depth, &total, &skipped);
        wr(out, &rc, 4);
        wr(out, accum, X * 16);
        wr(out, stats, sizeof stats);
        wr(out, counts, (size_t)n * 4);
        wr(out, offsets, (size_t)n * 4);
        wr(out, depth, S * 8);
        wr(out, &total, 4);
        wr(out, &skipped, 4);
        free(accum), free(depth), free(counts), free(offsets);
    }
    {   /* the sorted alpha path with the first parameter set */
        const uint32_t spp = *(const uint32_t *)params;
        float *accum = calloc(X * 4, 4);
        uint64_t stats[16] = {0};
        uint32_t *keys = calloc(n ? n : 1, 4), *vals = calloc(n ? n : 1, 4), vis = 0;
        int rc = mir_render_sorted(splats, n, rest ? sh : NULL, rest ? deg : 0, cam, params, W, H, spp, accum, stats,
                                   keys, vals, &vis);
        wr(out, &rc, 4);
        wr(out, accum, X * 16);
        wr(out, stats, sizeof stats);
        wr(out, keys, (size_t)n * 4);
        wr(out, vals, (size_t)n * 4);
        wr(out, &vis, 4);
        uint8_t *rgba = calloc(4 * X * 4, 1);
        mir_composite(accum, W, H, 1.0f, 1, 2 * W, 2 * H, rgba);
        wr(out, rgba, 4 * X * 4);
        free(accum), free(keys), free(vals), free(rgba);
    }
    {   /* the f64 renderers */
        const double bg[3] = {0.05, 0.05, 0.08};
        double *img = calloc(X * 3, 8), st[3];
        orc_render_particles_3d(orc, n, cam, W, H, 2, count, bg, seed, po, img, st);
        wr(out, img, X * 24);
        wr(out, st, sizeof st);
        orc_render_monte_carlo(orc, n, cam, W, H, 2, count, bg, seed, mo, img, st);
        wr(out, img, X * 24);
        wr(out, st, sizeof st);
        orc_render_analytic(orc, n, cam, W, H, bg, img);
        wr(out, img, X * 24);
        free(img);
    }
    fclose(out);
    free(hd), free(splats), free(sh), free(cam), free(params), free(orc), free(po), free(mo), free(lut);
    return 0;
}
