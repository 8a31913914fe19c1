/*
 * gps_hip: C ABI of the MI355X (gfx950) particle renderer.
 *
 * This ABI replaces the reference renderer gps_render::gps::GpsRenderer
 * (crates/gps-render/src/gps.rs:302-1065) for callers that bind C: the
 * gps-verify CLI (crates/gps-verify/src/main.rs), the phase gates
 * (crates/gps-render/src/verify.rs, lib.rs) and the browser viewer
 * (crates/gps-web/src/lib.rs).  Every method keeps the reference's name,
 * argument meaning and error behaviour:
 *
 *   - every entry point returns an int status: 0 = ok, nonzero = error, and
 *     gps_renderer_last_error() returns the message (the reference's
 *     Result<_, String>); a failed call leaves the renderer's state as it was
 *     (gps.rs:682);
 *   - all GPU work of one renderer runs in order on one HIP stream owned by
 *     the renderer; read functions synchronise that stream;
 *   - one renderer per host thread (the reference's &mut self).
 *
 * Struct layouts are the reference's (gps.rs:26-59, 167-272; oracle.rs:33-42).
 */
#ifndef GPS_HIP_H
#define GPS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One Gaussian, 64 bytes: GpuGaussian, gps.rs:26-38 */
typedef struct GpsGaussian {
    float pos_opacity[4]; /* xyz position, w = activated opacity */
    float log_scale[4];   /* xyz log standard deviations */
    float rot_wxyz[4];    /* quaternion (w, x, y, z), normalised by the kernels */
    float color[4];       /* xyz = f_dc * SH_C0 + 0.5 (linear) */
} GpsGaussian;

/* OracleCamera, oracle.rs:33-42.  cam = view_rot * (world - view_pos), +Z forward,
 * pixel y down; view_rot is row-major.  cx = cy = 0 means the image centre. */
typedef struct GpsCamera {
    double view_rot[9];
    double view_pos[3];
    double focal_x, focal_y, cx, cy, low_pass;
} GpsCamera;

/* Method, gps.rs:104-140 */
enum { GPS_METHOD_GPS = 0, GPS_METHOD_PROPORTIONAL = 1, GPS_METHOD_EXTINCTION = 2,
       GPS_METHOD_VIEW_CONDITIONED = 3, GPS_METHOD_SCREEN_OCCUPANCY = 4 };
/* Calibration, gps.rs:144-154 */
enum { GPS_CAL_NONE = 0, GPS_CAL_OBJECT_ZOOM = 1, GPS_CAL_PER_SPLAT_DEPTH = 2, GPS_CAL_PER_SPLAT_FOOTPRINT = 3 };
/* RenderPath, gps.rs:93-100.  GPS_PATH_PARTICLE_SORTED has no counterpart in the reference: particles are
 * resolved by a stable 64-bit sort of (ordered depth key << 32 | particle index) and the first hit per
 * subpixel, which gives the same depth|id buffers and images as GPS_PATH_PARTICLE bit for bit. */
enum { GPS_PATH_PRIMITIVE = 0, GPS_PATH_PARTICLE = 1, GPS_PATH_PARTICLE_SORTED = 2 };

/* PbvrParams, gps.rs:169-179 */
typedef struct GpsPbvrParams {
    uint32_t calibration;
    uint32_t radial_correction;
    uint32_t centre_depth;
    uint32_t low_pass_jitter;
    float base_k;
    float reference_pixel_length;
} GpsPbvrParams;

/* RenderParams, gps.rs:189-210 */
typedef struct GpsRenderParams {
    uint32_t spp_side;             /* subpixels per pixel = spp_side^2, 1..4 */
    float density_scale;
    uint32_t max_points_per_splat;
    float near;
    float background[3];
    float exposure;
    uint32_t srgb_output;
    uint32_t exact_radius;
    uint32_t sh_degree;
    uint32_t path;
    uint32_t method;
    GpsPbvrParams pbvr;
} GpsRenderParams;

/* GpuStats, gps.rs:233-252 (cumulative since the last gps_renderer_reset_stats) */
typedef struct GpsStats {
    uint64_t points;
    uint64_t truncated_splats;
    uint64_t dropped_points;
    uint64_t visible_splats;
    uint64_t orphan_subpixels; /* always 0: one packed 64-bit minimum per subpixel */
    uint64_t candidates;
    uint64_t undercovered_splats;
    uint64_t skipped_ensembles;
    uint64_t ensembles;
    uint64_t probe_demand;
    uint32_t issue;
    uint32_t reserved;
} GpsStats;

typedef struct GpsRenderer GpsRenderer;

/* RenderParams::default(), gps.rs:212-229 */
void gps_render_params_default(GpsRenderParams *out);

/* GpsRenderer::new, gps.rs:380-570: w x h pixels, spp_side clamped to 1..4, on HIP device `device`. */
int gps_renderer_create(uint32_t width, uint32_t height, uint32_t spp_side, int device, GpsRenderer **out);
void gps_renderer_destroy(GpsRenderer *r);
/* message of the last failed call on r (empty if none); with r == NULL, of the last failed create */
const char *gps_renderer_last_error(const GpsRenderer *r);

/* set_gaussians / set_gaussians_sh, gps.rs:683-740.  sh_rest: NULL or 3 * ((degree+1)^2 - 1) floats per
 * splat, channel-major; degree 1..3. */
int gps_renderer_set_gaussians(GpsRenderer *r, const GpsGaussian *g, uint64_t n, const float *sh_rest,
                               uint32_t degree);
/* resize, gps.rs:781-789 */
int gps_renderer_resize(GpsRenderer *r, uint32_t width, uint32_t height, uint32_t spp_side);
/* reset_accum, gps.rs:792-801 (also restarts the skipped-ensemble counter) */
int gps_renderer_reset_accum(GpsRenderer *r);
/* reset_stats, gps.rs:803-809 */
int gps_renderer_reset_stats(GpsRenderer *r);
/* render_ensembles, gps.rs:861-925: ensembles first_seed .. first_seed+count-1 into the accumulation
 * buffer.  Any count is accepted (the reference caps one submission at 4). */
int gps_renderer_render_ensembles(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p,
                                  uint32_t first_seed, uint32_t count);
/* render_sorted, gps.rs:930-944 */
int gps_renderer_render_sorted(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p,
                               uint32_t as_ensembles);
/* probe_demand, gps.rs:948-955 */
int gps_renderer_probe_demand(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p);
/* composite, gps.rs:967-986, into an out_w x out_h RGBA8 image (the wgpu target view's replacement) */
int gps_renderer_composite(GpsRenderer *r, const GpsRenderParams *p, uint32_t out_w, uint32_t out_h,
                           uint8_t *rgba8);
/* read_accum, gps.rs:989-993: w*h*4 floats (RGB sum, ensemble count) */
int gps_renderer_read_accum(GpsRenderer *r, float *out);
/* read_stats, gps.rs:996-1000 */
int gps_renderer_read_stats(GpsRenderer *r, GpsStats *out);
/* read_sorted, gps.rs:958-963: n keys and n values of the last render_sorted */
int gps_renderer_read_sorted(GpsRenderer *r, uint32_t *keys, uint32_t *vals);
/* request_stats, gps.rs:1005-1027: starts a non-blocking readback of the counters on the renderer's stream, after
 * the work enqueued so far; a no-op while one is in flight.  Each readback gets the next issue number. */
int gps_renderer_request_stats(GpsRenderer *r);
/* latest_stats, gps.rs:1035-1038: the most recent counters a readback delivered, with its issue number in
 * `issue` (all zero before the first); delivers a finished readback first, never waits */
int gps_renderer_latest_stats(GpsRenderer *r, GpsStats *out);
/* stats_issued, gps.rs:1029-1033: readbacks issued so far.  A snapshot whose issue is greater than the value read
 * right after some work was enqueued was copied after that work completed. */
uint32_t gps_renderer_stats_issued(const GpsRenderer *r);
/* wait_idle, gps.rs:1055-1057 */
int gps_renderer_wait_idle(GpsRenderer *r);
/* submitted / accumulated, gps.rs:669-677 */
uint32_t gps_renderer_submitted(const GpsRenderer *r);
uint32_t gps_renderer_accumulated(GpsRenderer *r);
/* size / spp_side / gaussian_count / sh_max_degree accessors, gps.rs:651-666 */
void gps_renderer_info(const GpsRenderer *r, uint32_t *width, uint32_t *height, uint32_t *spp_side,
                       uint64_t *n_splats, uint32_t *sh_max_degree);

/* estimate_memory, gps.rs:288-291 (bytes of depth+winner, accum, splats) */
void gps_estimate_memory(uint32_t width, uint32_t height, uint32_t spp_side, uint64_t n_splats,
                         uint64_t *depth_winner, uint64_t *accum, uint64_t *splats);
/* radius_lut, crates/gps-core/src/lib.rs:624-636: 65 x 129 floats */
void gps_radius_lut(float *out);

/* ---- bit-exact test hooks (the reference reads these buffers in its phase-0 and verify gates) ----
 * GPS_DEBUG_KEEP_BUFFERS keeps the per-ensemble buffers of the last ensemble of the last
 * gps_renderer_render_ensembles call readable (one ensemble per batch; slower). */
enum { GPS_DEBUG_KEEP_BUFFERS = 1 };
int gps_renderer_set_debug(GpsRenderer *r, uint32_t flags);
/* per-splat particle counts and exclusive offsets (n u32 each), particle total and skip flag */
int gps_renderer_read_counts(GpsRenderer *r, uint32_t *counts, uint32_t *offsets, uint32_t *total,
                             uint32_t *skipped);
/* packed depth|id per subpixel (w*h*spp u64): (ordered depth key << 32) | splat id, all ones when empty;
 * the high and low halves are the reference's `depth` and `winner` buffers (gps_splat.wgsl:28-29) */
int gps_renderer_read_depth_id(GpsRenderer *r, uint64_t *out);

/* Particle kernel counters since the last reset_stats (no counterpart in the reference; measurement and test
 * hooks): out[0] the particles that reached a subpixel (on screen: a packed-minimum update or an on-screen sorted
 * key), the HBM-writing part of the particle work; the high-water marks of the count-only LDS buffers against
 * their capacities: out[1] a wave's emit_pbvr queue's peak fill, out[2] its capacity (128), out[3] a wave's
 * count_only work buffer's peak fill, out[4] its capacity (128). */
int gps_renderer_read_particle_counters(GpsRenderer *r, uint64_t *out);

/* GPS_PATH_PARTICLE_SORTED: the sorted keys and subpixels (0xFFFFFFFF = off screen) of the last ensemble;
 * copies min(count, max) entries and stores the particle count in *count. */
int gps_renderer_read_particle_keys(GpsRenderer *r, uint64_t *keys, uint32_t *subpixels, uint64_t max,
                                    uint64_t *count);

/* Stable LSD radix sort of n (u64 key, u32 value) pairs in place (the onesweep sort of the sorted path);
 * n must be below 2^32. */
int gps_sort_pairs_u64(int device, uint64_t *keys, uint32_t *vals, uint64_t n);

/* Phase-0 unit kernels (crates/gps-render/src/lib.rs:99-275; shaders/p0_sample.wgsl, gps_scan.wgsl).
 * gps_phase0_atomic: samples of p0_sample.wgsl:205-218 for one ensemble with one packed 64-bit
 * minimum; writes pixels depth keys and colours.  gps_scan_exclusive_u32: exclusive prefix sum. */
int gps_phase0_atomic(int device, uint32_t seed, uint32_t pixels, uint32_t samples, uint32_t *depth,
                      uint32_t *color);
int gps_scan_exclusive_u32(int device, const uint32_t *in, uint32_t *out, uint64_t n);
/* Allocation-free variant of gps_scan_exclusive_u32 (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  `in`
 * and `out` are device buffers of n u32 on `device` (they may alias); the scan is enqueued on `stream` (a hipStream_t
 * of that device, NULL for the null stream) and not waited for.  It calls no hipMalloc, hipFree or allocating
 * routine: the look-back scratch (tile ticket and status words, 1 MiB + 256 B) is a file-scope device array that
 * each GPU holds until the process ends, cleared by a fill kernel at every call.  Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (on different streams or host threads) race.  Makes `device`
 * current.  n above GPS_SCAN_NOALLOC_MAX_N returns GPS_SCAN_NOALLOC_TOO_LARGE and enqueues nothing. */
#define GPS_SCAN_NOALLOC_MAX_N (1ull << 28)
enum { GPS_SCAN_NOALLOC_TOO_LARGE = -2 };
int gps_scan_exclusive_u32_noalloc(int device, const uint32_t *in, uint32_t *out, uint64_t n, void *stream);
/* Device evaluation of the kernels' f32 functions (0 exp, 1 log, 2 sin, 3 cos, 4 sqrt, 5 div x/y). */
int gps_math_eval(int device, int fn, const float *x, const float *y, float *out, uint64_t n);

/* ---- denoiser features (no counterpart in the reference).  When enabled, every resolve also accumulates per
 * pixel the sum over ensembles of the squared ensemble pixel value (r, g, b) and the number of covered
 * subpixels (w), and keeps the nearest particle depth seen (+inf where nothing was hit), in
 * gps_renderer_render_ensembles and batches on every particle path (gps_renderer_render_sorted's alpha path
 * does not resolve subpixels and leaves them alone).  reset_accum clears them with the accumulation buffer.
 * Enabling clears them; disabling frees them. */
int gps_renderer_set_features(GpsRenderer *r, int enable);
/* Skips PBVR splats (Proportional, Extinction, ViewConditioned) none of whose particles can land in the
 * image: a conservative bound from the clamped normal draws (no counterpart in the reference, which
 * generates and then discards them).  Images and depth|id buffers are bit-identical with and without it;
 * the points, candidates and visible_splats counters then leave those splats out.  Off by default. */
int gps_renderer_set_cull(GpsRenderer *r, int enable);
/* moments: w*h*4 floats (may be NULL); zmin: w*h floats (may be NULL) */
int gps_renderer_read_features(GpsRenderer *r, float *moments, float *zmin);

/* ---- learned denoiser (no counterpart in the reference).  A 5-layer 3x3 convolution network (32 channels,
 * ReLU, residual output) maps 7 per-pixel features of a renderer's accumulation (mean rgb, per-ensemble standard
 * deviation, standard error of the mean, 1/(1 + nearest depth), covered fraction) to the many-ensemble image.
 * BF16 weights and activations on MFMA with FP32 accumulation; FP32 master weights trained with Adam.
 * Parameter vector: per layer l = 0..4, W[cout][3 x 3 taps][cin] then b[cout], with (cin, cout) =
 * (8, 32), (32, 32), (32, 32), (32, 32), (32, 3) (input channel 7 is zero).  A tile of 16 x 16 pixels falls back
 * to the plain mean when the mean squared change the network makes exceeds fallback_ratio times the variance
 * of the mean (estimated from the moments); fallback_ratio <= 0 turns the fallback off. */
typedef struct GpsDenoiser GpsDenoiser;
typedef struct GpsDenoiseInfo {
    double gpu_ms; /* features + network + fallback, HIP events */
    uint32_t width, height, tiles, fallback_tiles;
} GpsDenoiseInfo;
uint64_t gps_denoiser_param_count(void);
int gps_denoiser_create(int device, uint32_t seed, GpsDenoiser **out);
void gps_denoiser_destroy(GpsDenoiser *d);
const char *gps_denoiser_last_error(const GpsDenoiser *d);
int gps_denoiser_set_params(GpsDenoiser *d, const float *params); /* also resets Adam */
int gps_denoiser_get_params(GpsDenoiser *d, float *params);
/* denoises renderer r's accumulation (its features must be enabled); waits for the renderer's work */
int gps_denoiser_apply(GpsDenoiser *d, GpsRenderer *r, float fallback_ratio, GpsDenoiseInfo *info);
/* apply's forward: 1 (default) one fused kernel for the five layers, whose intermediate activations stay on chip;
 * 0 the layer-by-layer kernels the trainer runs.  Both give the same bits.  Returns 1 when this build's network
 * has no fused kernel (it then always runs layer by layer). */
int gps_denoiser_set_fused(GpsDenoiser *d, int enable);
int gps_denoiser_read(GpsDenoiser *d, float *rgba);                        /* w*h*4 of the last apply */
int gps_denoiser_read_features(GpsDenoiser *d, float *x8, float *mean4);    /* last apply's inputs (w*h*8, w*h*4) */
/* training pairs: the input renderer's features and the target renderer's mean (same size for every pair) */
int gps_denoiser_add_pair(GpsDenoiser *d, GpsRenderer *input, GpsRenderer *target); /* returns the pair index */
uint32_t gps_denoiser_pairs(const GpsDenoiser *d);
/* per pair: sqrt(the input's ensemble count / 4); the training loss weights each pair's squared error by it */
int gps_denoiser_pair_weights(GpsDenoiser *d, float *out);
int gps_denoiser_set_pair_weights(GpsDenoiser *d, const float *weights); /* one per pair */
int gps_denoiser_clear_pairs(GpsDenoiser *d);
/* `steps` Adam steps on batches of `batch` pairs drawn with a PCG32 stream seeded by `seed`; losses[steps]
 * (the pair-weighted mean squared error over the batch's pixels and channels) may be NULL */
int gps_denoiser_train(GpsDenoiser *d, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float *losses);
/* loss and parameter gradient over pairs 0..batch-1 without an update (test hook) */
int gps_denoiser_loss_grad(GpsDenoiser *d, uint32_t batch, float *loss, float *grad);
int gps_denoiser_read_output(GpsDenoiser *d, uint32_t images, float *rgba); /* network output of the last step */

/* ---- batches of views (no counterpart in the reference: its trajectory export renders one view per frame loop,
 * web/index.html:445-469).  Each item renders `count` ensembles of one view (a renderer's scene, camera and
 * parameters) from an empty image and zero counters into the batch's slot `i`; afterwards each renderer holds
 * the image, counters and ensemble count of its last item, and continues that history exactly.
 *
 * Fused (the default for items on the particle path of non-empty scenes without the debug hook, up to 1024
 * items): every item gets buffers of its own in the batch, and the whole batch runs through one set of
 * kernels: a clear, one view kernel for all items, then per round of ensembles (the tuning's batch_ensembles,
 * or as many as keep the largest item's depth|id buffers within depth_window_mib) one count-and-scan launch over
 * every (item, ensemble), per emitter class one assignment of the particle grid to the items in proportion to
 * their particles and one particle launch, and one resolve launch over every item's pixels.  With
 * GPS_BATCH_GRAPH these launches are captured into one hipGraph (a few nodes per round) and launched once.
 *
 * GPS_BATCH_PER_ITEM records each item's own launches instead (the single-view kernels), items of different
 * renderers on up to four forked streams (parallel branches of the graph); it is also taken when an item
 * cannot be fused.  Results are bit-identical either way.  The sorted particle path is not batchable. */
enum { GPS_BATCH_GRAPH = 1, GPS_BATCH_PER_ITEM = 2 };
typedef struct GpsBatch GpsBatch;
typedef struct GpsBatchItem {
    GpsRenderer *renderer;
    const GpsCamera *camera;
    const GpsRenderParams *params;
    uint32_t first_seed;
    uint32_t count;
} GpsBatchItem;
int gps_batch_create(int device, GpsBatch **out);
void gps_batch_destroy(GpsBatch *b);
/* flags: GPS_BATCH_GRAPH and / or GPS_BATCH_PER_ITEM (0: the fused launches straight onto the batch's stream) */
int gps_batch_render(GpsBatch *b, const GpsBatchItem *items, uint32_t n, int flags);
/* w*h*4 floats of item `item`'s accumulation snapshot */
int gps_batch_read(GpsBatch *b, uint32_t item, float *out);
/* item `item`'s counters: every item also resets its renderer's counters before rendering */
int gps_batch_read_stats(GpsBatch *b, uint32_t item, GpsStats *out);
/* composite (gps.rs:967-986, as gps_renderer_composite) of item `item`'s snapshot into out_w*out_h RGBA8 */
int gps_batch_composite(GpsBatch *b, uint32_t item, const GpsRenderParams *p, uint32_t out_w, uint32_t out_h,
                        uint8_t *rgba8);
/* GPU time of the last gps_batch_render (graph launch, or the fused launches, to completion; the host's time for
 * per-item launches without a graph), host time to record it, graph nodes */
void gps_batch_timing(const GpsBatch *b, double *gpu_ms, double *record_ms, uint64_t *graph_nodes);

/* ---- 3DGS PLY ingest (crates/gps-io/src/lib.rs:1-373; the activation of crates/gps-render/src/gps.rs:41-49).
 * ASCII or binary little-endian, one vertex element first, scalar properties of any PLY type, SH degree from
 * 0 / 9 / 24 / 45 f_rest_* properties; other properties are skipped.  Values are kept in file form
 * (log-scales, opacity logit, (w, x, y, z) rotation, f_dc, channel-major f_rest) until gps_ply_to_gpu.
 * Errors: a GPS_PLY_* code (the PlyError variants, lib.rs:16-26) and gps_ply_last_error()'s message. */
#define GPS_PLY_MAX_VERTICES 200000000ull /* lib.rs:13 */
enum {
    GPS_PLY_NOT_PLY = 1,
    GPS_PLY_BAD_HEADER = 2,
    GPS_PLY_UNSUPPORTED = 3,
    GPS_PLY_MISSING_PROPERTY = 4,
    GPS_PLY_TOO_MANY_VERTICES = 5,
    GPS_PLY_TRUNCATED = 6,
    GPS_PLY_BAD_ASCII = 7,
    GPS_PLY_NON_FINITE = 8,
    GPS_PLY_BAD_SH_COUNT = 9
};
typedef struct GpsPly GpsPly;
/* load_ply (lib.rs:260): parses `len` bytes; 0 and *out, or a GPS_PLY_* code */
int gps_ply_load(const uint8_t *bytes, uint64_t len, GpsPly **out);
void gps_ply_free(GpsPly *g);
const char *gps_ply_last_error(void); /* message of the calling thread's last failed load */
void gps_ply_info(const GpsPly *g, uint64_t *count, uint32_t *sh_degree, int *binary);
/* file-form arrays (any may be NULL): 3, 3, 4, 1, 3 and 3 * rest-per-channel floats per splat */
void gps_ply_read(const GpsPly *g, float *position, float *log_scale, float *rot_wxyz, float *opacity_logit,
                  float *f_dc, float *sh_rest);
/* GpuGaussian::from_activated(Gaussians::activated(i)) for every splat: opacity = sigmoid(logit) in f64,
 * colour = f_dc * SH_C0 + 0.5 in f64, the rest as stored; pass sh_rest from gps_ply_read with the degree to
 * gps_renderer_set_gaussians */
void gps_ply_to_gpu(const GpsPly *g, GpsGaussian *out);

/* ---- launch tuning (process-wide; no counterpart in the reference).  Rendered bits never depend on it:
 * it only changes how the work is split into launches.  Renderers pick it up on their next call. */
typedef struct GpsTuning {
    uint32_t particle_waves_per_cu;  /* particle grid = CUs x this workgroups (1..1024), default 64; grids
                                        larger than the resident workgroups balance uneven chunks better */
    uint32_t particles_per_thread;   /* 1, 2, 4 (default) or 8 */
    uint32_t sort_items_per_thread;  /* radix sort tile of 256 x this keys: 8, 16, 32 or 64 (default; run
                                        as 512 or 1024 threads x 16) */
    uint32_t batch_ensembles;        /* ensembles per particle launch (1..64); 0 = from depth_window_mib */
    uint32_t depth_window_mib;       /* with batch_ensembles 0: as many ensembles per launch as keep their
                                        depth|id buffers (8 bytes per subpixel each) within this many MiB,
                                        at least 1; default 64 */
    uint32_t count_only;             /* 1 (default): PBVR particles of splats whose particles provably all land
                                        off-screen are evaluated for the generated-points count on the certified
                                        count-only path (undecided particles on the full path); 0: every particle
                                        on the full path.  Images and counters are the same either way */
    uint32_t particle_block;         /* particle kernel workgroup size: 128, 256 (default) or 512 threads */
    uint32_t sort_digit_bits;        /* radix digit width of the 64-bit sort: 4..8 bits (default 8) */
} GpsTuning;
void gps_tuning_get(GpsTuning *out);
int gps_tuning_set(const GpsTuning *t);
/* Autotuner: times every candidate (workgroup size 128/256/512 x 16/32/64/128/256 256-thread groups' worth of
 * threads per CU x particles per thread 1/2/4/8, and sort tiles 8/16/32/64
 * on the sorted particle path) on trial_ensembles ensembles of this scene and view, then the depth window
 * (8 .. 4096 MiB, with batch_ensembles 0) on max(trial_ensembles, 2 x the ensembles per launch it gives),
 * keeps the fastest as the process tuning, and restores the accumulation buffer, denoiser features, counters
 * and submitted count; on the sorted particle path it then times the sort's digit widths 4..8 bits.  Writes the
 * chosen tuning and its GPU ms per ensemble. */
int gps_renderer_autotune(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p, uint32_t first_seed,
                          uint32_t trial_ensembles, GpsTuning *best, double *best_ms_per_ensemble);

/* ---- timing: HIP events around every kernel launch on the renderer's stream ---- */
typedef struct GpsKernelTime {
    char name[32];
    uint64_t launches;
    double total_ms;
} GpsKernelTime;
int gps_renderer_profile(GpsRenderer *r, int enable);
/* fills up to max entries, returns the number of kernels recorded (or -1) */
int gps_renderer_profile_read(GpsRenderer *r, GpsKernelTime *out, int max);

#ifdef __cplusplus
}
#endif
#endif
