// Host side of the renderer: the C ABI of include/gps_hip.h over the kernels of gps_kernels.hip.
//
// Mirrors gps_render::gps::GpsRenderer (crates/gps-render/src/gps.rs:302-1065): it owns every device
// buffer, packs the per-view constants like upload_frame (gps.rs:812-856), and runs ensembles in
// batches (one depth|id buffer per ensemble of a batch) on one in-order HIP stream.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gps_hip.h"
#include "gps_internal.h"

// sorted path (gps_sorted.hip)
struct SortedState;
SortedState *sorted_create();
void sorted_destroy(SortedState *s);
int sorted_render(SortedState *s, hipStream_t st, const GpsFrame &f, const float4 *splats, const float *sh,
                  float4 *accum, uint64_t *stats, int probe_only, std::string &err,
                  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> *prof);
int sorted_read(SortedState *s, hipStream_t st, uint32_t n, uint32_t *keys, uint32_t *vals, std::string &err);

// 64-bit onesweep sort and the depth-sorted resolve (gps_onesweep.hip)
struct OnesweepState;
OnesweepState *onesweep_create();
void onesweep_destroy(OnesweepState *s);
int onesweep_sort(OnesweepState *s, hipStream_t st, uint64_t *k[2], uint32_t *v[2], uint64_t n, int &result,
                  std::string &err, std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> *prof,
                  int lo_bit);
void launch_first_hit(hipStream_t s, const uint32_t *vals, uint64_t n, uint32_t *first, uint64_t S);
void launch_build_depth(hipStream_t s, const uint64_t *keys, const uint32_t *first, uint64_t S,
                        const uint32_t *offsets, uint32_t N, uint64_t *depth);

namespace {

thread_local std::string g_create_error;
GpsTuningDev g_tuning = {64u, 4u, 64u, 0u, 64u, 1u, 256u, 8u};

constexpr uint64_t MAX_SPLATS = 1ull << 30;
constexpr uint32_t LUT_O = 65, LUT_V = 129;

// ------------------------------------------------------------------ radius table (f64, lib.rs:178-248, 522-636)
double dilog_series(double x) {
    double term = x, sum = x;
    for (int k = 2; k < 1000; k++) {
        term *= x;
        double add = term / ((double)k * (double)k);
        sum += add;
        if (std::fabs(add) < 1e-18 * std::fabs(sum)) break;
    }
    return sum;
}

constexpr double DILOG_AT_ONE = 1.6449340668482264;

double dilog64(double x) {
    if (x >= 1.0) return DILOG_AT_ONE;
    if (x <= -1.0) return -DILOG_AT_ONE / 2.0;
    if (x == 0.0) return 0.0;
    if (x < 0.0) {
        double y = x / (x - 1.0);
        double l = std::log(1.0 - x);
        return -dilog_series(y) - 0.5 * l * l;
    }
    if (x <= 0.5) return dilog_series(x);
    return DILOG_AT_ONE - std::log(x) * std::log(1.0 - x) - dilog_series(1.0 - x);
}

double inv_dilog64(double target) {
    double below_one = std::nextafter(DILOG_AT_ONE, 0.0);
    target = target < 0.0 ? 0.0 : (target > below_one ? below_one : target);
    if (target <= 0.0) return 0.0;
    double lo = 0.0, hi = 1.0, y = target / DILOG_AT_ONE;
    for (int i = 0; i < 40; i++) {
        double f = dilog64(y) - target;
        if (f > 0.0) hi = y; else lo = y;
        double dfdy = (y > 1e-12 && y < 1.0 - 1e-15) ? -std::log(1.0 - y) / y : 1.0;
        double ny = y - f / dfdy;
        if (!(ny > lo && ny < hi)) ny = 0.5 * (lo + hi);
        if (std::fabs(ny - y) <= 1e-14 * (1.0 + y)) return ny;
        y = ny;
    }
    return y;
}

double corrected_radius64(double opacity, double u) {
    double o = opacity < 0.0 ? 0.0 : (opacity > 1.0 ? 1.0 : opacity);
    if (o <= 0.0) return 0.0;
    u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    double d = dilog64(o);
    double y = std::fmin(inv_dilog64((1.0 - u) * d), o);
    if (y <= 0.0) return 40.0;
    return std::sqrt(std::fmax(-2.0 * std::log(y / o), 0.0));
}

void make_lut(float *out) {
    for (uint32_t i = 0; i < LUT_O; i++) {
        double x = (double)i / (double)(LUT_O - 1);
        double o = 1.0 - (1.0 - x) * (1.0 - x);
        o = o < 1e-6 ? 1e-6 : (o > 1.0 ? 1.0 : o);
        for (uint32_t j = 0; j < LUT_V; j++) {
            double w = (double)j * std::sqrt(17.0) / (double)(LUT_V - 1);
            double u = -std::expm1(-w * w);
            out[i * LUT_V + j] = (float)corrected_radius64(o, u);
        }
    }
}

}  // namespace

GpsTuningDev gps_tuning_current() { return g_tuning; }

struct GpsRenderer {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t W = 0, H = 0, spp_side = 1;
    uint64_t n = 0;
    uint32_t sh_max_degree = 0, sh_stride = 0;
    double centroid[3] = {0, 0, 0};
    std::string err;

    // scene
    float4 *splats = nullptr;
    float *sh = nullptr;
    float *lut = nullptr;
    float4 *view = nullptr;
    uint32_t *perm = nullptr;   // [n + 3]: count-only splat order (marked splats first), then its counter word
    uint2 *cinfo = nullptr;
    float4 *colour = nullptr;
    // batch
    uint32_t *offsets = nullptr;   // [batch_cap][n]
    uint64_t *depth = nullptr;     // [batch_cap][S]
    uint32_t batch_cap = 0;
    uint32_t tiles = 0;
    BatchState bs{};
    uint32_t epoch = 1;
    bool depth_dirty = false;
    // frame
    float4 *accum = nullptr;
    float4 *mom = nullptr;  // denoiser features (set_features): sum of squares and covered subpixels
    float *zmin = nullptr;  // nearest depth per pixel
    bool features = false;
    bool cull = false;      // skip PBVR splats whose particles all miss the image (set_cull)
    uint64_t *stats = nullptr;
    // non-blocking stats readback (request_stats / latest_stats / stats_issued): a pinned host copy of the
    // counters, the event that marks its copy done, the issue number it carries, and the last delivered snapshot
    uint64_t *stats_staging = nullptr;
    hipEvent_t stats_event = nullptr;
    bool stats_pending = false;
    uint32_t stats_issued = 0, stats_pending_issue = 0;
    GpsStats stats_latest{};
    uint32_t *composite_buf = nullptr;
    uint64_t composite_cap = 0;
    uint32_t accumulated = 0;
    uint32_t debug = 0;
    // last particle-path batch slot for the debug hooks
    bool have_last = false;
    SortedState *sorted = nullptr;
    // depth-sorted particle resolve (GPS_PATH_PARTICLE_SORTED)
    OnesweepState *os = nullptr;
    uint64_t *pkeys[2] = {nullptr, nullptr};
    uint32_t *pidx[2] = {nullptr, nullptr};
    uint64_t pcap = 0;
    uint32_t *first = nullptr;
    uint64_t first_cap = 0;
    int psort_result = 0;
    uint64_t last_np = 0;
    // the stream render work is enqueued on: the renderer's own, or a batch stream under graph capture
    hipStream_t enqueue_stream = nullptr;
    // profiling
    bool profiling = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::map<std::string, std::pair<uint64_t, double>> times;
};

namespace {

int fail(GpsRenderer *r, const std::string &m) {
    r->err = m;
    return -1;
}

#define HIP_TRY(r, call)                                                                     \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((r), std::string(#call) + ": " + hipGetErrorString(e_));             \
    } while (0)

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

uint64_t sub_count(const GpsRenderer *r) { return (uint64_t)r->W * r->H * r->spp_side * r->spp_side; }

// Per-launch HIP events on the renderer's stream when profiling is on (never inside a batch graph).
struct Prof {
    GpsRenderer *r;
    const char *name;
    bool on;
    hipEvent_t a = nullptr, b = nullptr;
    Prof(GpsRenderer *r_, const char *n_) : r(r_), name(n_), on(r_->profiling && r_->enqueue_stream == r_->stream) {
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, r->stream);
        }
    }
    ~Prof() {
        if (on) {
            (void)hipEventRecord(b, r->stream);
            r->pending.push_back({name, {a, b}});
        }
    }
};

void drain_profile(GpsRenderer *r) {
    for (auto &p : r->pending) {
        float ms = 0.0f;
        (void)hipEventSynchronize(p.second.second);
        (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
        auto &t = r->times[p.first];
        t.first += 1;
        t.second += ms;
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    r->pending.clear();
}

int alloc_frame(GpsRenderer *r) {
    dfree(r->accum);
    dfree(r->depth);
    dfree(r->mom);
    dfree(r->zmin);
    uint64_t X = (uint64_t)r->W * r->H;
    HIP_TRY(r, hipMalloc(&r->accum, X * sizeof(float4)));
    launch_accum_clear(r->stream, r->accum, X);
    if (r->features) {
        HIP_TRY(r, hipMalloc(&r->mom, X * sizeof(float4)));
        HIP_TRY(r, hipMalloc(&r->zmin, X * sizeof(float)));
        launch_features_clear(r->stream, r->mom, r->zmin, X);
    }
    r->accumulated = 0;
    return 0;
}

// depth|id buffers and per-ensemble offsets sized for a batch of ensembles
int alloc_batch(GpsRenderer *r) {
    dfree(r->depth);
    dfree(r->offsets);
    dfree(r->bs.status);
    uint64_t S = sub_count(r);
    const char *env = std::getenv("GPS_HIP_BATCH_BYTES");
    uint64_t budget = env ? std::strtoull(env, nullptr, 10) : (8ull << 30);
    uint64_t per = 8 * S + 4 * (r->n ? r->n : 1);
    uint64_t cap = budget / per;
    if (cap < 1) cap = 1;
    if (cap > MAX_BATCH) cap = MAX_BATCH;
    r->batch_cap = (uint32_t)cap;
    HIP_TRY(r, hipMalloc(&r->depth, cap * S * sizeof(uint64_t)));
    launch_fill_u64(r->stream, r->depth, cap * S, EMPTY64);
    HIP_TRY(r, hipMalloc(&r->offsets, cap * (r->n ? r->n : 1) * sizeof(uint32_t)));
    r->tiles = (uint32_t)((r->n + SCAN_TILE - 1) / SCAN_TILE);
    if (r->tiles == 0) r->tiles = 1;
    HIP_TRY(r, hipMalloc(&r->bs.status, (uint64_t)MAX_BATCH * r->tiles * sizeof(uint64_t)));
    HIP_TRY(r, hipMemsetAsync(r->bs.status, 0, (uint64_t)MAX_BATCH * r->tiles * sizeof(uint64_t), r->stream));
    r->depth_dirty = false;
    r->have_last = false;
    return 0;
}

void pack_frame(const GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p, GpsFrame &f) {
    std::memset(&f, 0, sizeof f);
    for (int row = 0; row < 3; row++) {
        for (int k = 0; k < 3; k++) f.rot[row][k] = (float)cam->view_rot[3 * row + k];
        f.rot[row][3] = 0.0f;
    }
    for (int k = 0; k < 3; k++) f.pos[k] = (float)cam->view_pos[k];
    f.pos[3] = (float)cam->low_pass;
    f.jitter = std::sqrt(f.pos[3] > 0.0f ? f.pos[3] : 0.0f);   // IEEE f32 sqrt, the kernels' value bit for bit
    double cx = cam->cx, cy = cam->cy;
    if (cx == 0.0 && cy == 0.0) {
        cx = 0.5 * (double)r->W;
        cy = 0.5 * (double)r->H;
    }
    f.intr[0] = (float)cam->focal_x;
    f.intr[1] = (float)cam->focal_y;
    f.intr[2] = (float)cx;
    f.intr[3] = (float)cy;
    f.W = r->W;
    f.H = r->H;
    f.spp_side = r->spp_side;
    f.N = (uint32_t)r->n;
    f.cap = p->max_points_per_splat > 1u ? p->max_points_per_splat : 1u;
    f.sh_stride = r->sh_stride;
    f.exact_radius = p->exact_radius ? 1u : 0u;
    f.sh_degree = p->sh_degree < r->sh_max_degree ? p->sh_degree : r->sh_max_degree;
    f.method = p->method;
    f.calib = p->pbvr.calibration;
    uint32_t flags = 0;
    if (p->pbvr.radial_correction) flags |= FLAG_RADIAL;
    if (p->pbvr.centre_depth) flags |= FLAG_CENTRE;
    if (p->pbvr.low_pass_jitter) flags |= FLAG_JITTER;
    if (r->cull) flags |= FLAG_CULL;
    const bool pbvr = p->method == M_PROPORTIONAL || p->method == M_EXTINCTION || p->method == M_VIEW_CONDITIONED;
    // the count-only records carry (splat, draw chain), not a particle index, so the path holds for every cap
    if (g_tuning.count_only && pbvr) flags |= FLAG_FAST;
    f.flags = flags;
    f.density = p->density_scale;
    f.near_z = p->near;
    f.base_k = p->pbvr.base_k;
    f.ref_len = p->pbvr.reference_pixel_length;
    double d[3] = {r->centroid[0] - cam->view_pos[0], r->centroid[1] - cam->view_pos[1],
                   r->centroid[2] - cam->view_pos[2]};
    f.object_depth = (float)(cam->view_rot[6] * d[0] + cam->view_rot[7] * d[1] + cam->view_rot[8] * d[2]);
    for (int k = 0; k < 3; k++) f.bg[k] = p->background[k];
    f.wf = (float)f.W;
    f.hf = (float)f.H;
    f.sidef = (float)f.spp_side;
}

int check_params(GpsRenderer *r, const GpsRenderParams *p) {
    uint32_t s = p->spp_side < 1 ? 1 : (p->spp_side > 4 ? 4 : p->spp_side);
    if (s != r->spp_side) return fail(r, "params.spp_side differs from the renderer's; call resize() first");
    if (p->method > 4) return fail(r, "unknown method");
    if (p->path > GPS_PATH_PARTICLE_SORTED) return fail(r, "unknown render path");
    return 0;
}

// Ensembles per particle launch: the tuned count, else as many as keep the launch's depth|id buffers within
// the tuned window (the packed minimum's read-modify-writes then stay in the last-level cache), at most the
// allocated batch.
uint32_t ensembles_per_launch(const GpsRenderer *r) {
    uint32_t cap = r->batch_cap;
    if (g_tuning.batch_ensembles) {
        if (cap > g_tuning.batch_ensembles) cap = g_tuning.batch_ensembles;
        return cap;
    }
    const uint64_t window = (uint64_t)g_tuning.depth_window_mib << 20;
    const uint64_t per = 8 * sub_count(r);
    uint64_t b = window / (per ? per : 1);
    if (b < 1) b = 1;
    if (b < cap) cap = (uint32_t)b;
    return cap;
}

int grid_for_particles(GpsRenderer *r) {
    static int cus = 0;
    if (!cus) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, r->device) == hipSuccess) cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    int per_cu = (int)g_tuning.particle_waves_per_cu;
    if (per_cu < 1) per_cu = 1;
    return cus * per_cu;
}

// One ensemble at a time: counts + scan, per-particle keys, onesweep sort, first hit per subpixel,
// depth|id from the first hit, resolve.  The particle total is read back to size the sort.
int render_sorted_particles(GpsRenderer *r, const GpsFrame &f, uint32_t first_seed, uint32_t count) {
    r->enqueue_stream = r->stream;
    const uint64_t S = sub_count(r);
    const bool keep = (r->debug & GPS_DEBUG_KEEP_BUFFERS) != 0;
    const int grid = grid_for_particles(r);
    if (r->first_cap < S) {
        dfree(r->first);
        HIP_TRY(r, hipMalloc(&r->first, S * 4));
        r->first_cap = S;
    }
    for (uint32_t e = 0; e < count; e++) {
        const uint32_t seed = first_seed + e;
        if (r->depth_dirty) {
            launch_fill_u64(r->stream, r->depth, S, EMPTY64);
            r->depth_dirty = false;
        }
        {
            Prof pr(r, "count_scan");
            launch_count_scan(r->stream, f, r->cinfo, seed, 1, r->offsets, r->bs, r->tiles, r->epoch, r->stats);
        }
        r->epoch = (r->epoch + 1u) & ((1u << 26) - 1u);
        if (r->epoch == 0) r->epoch = 1;
        uint32_t total = 0;
        HIP_TRY(r, hipMemcpyAsync(&total, r->bs.totals, 4, hipMemcpyDeviceToHost, r->stream));
        HIP_TRY(r, hipStreamSynchronize(r->stream));
        if (total > 0) {
            if (total > r->pcap) {
                for (int k = 0; k < 2; k++) {
                    dfree(r->pkeys[k]);
                    dfree(r->pidx[k]);
                }
                uint64_t cap = (uint64_t)total + total / 4;
                for (int k = 0; k < 2; k++) {
                    HIP_TRY(r, hipMalloc(&r->pkeys[k], cap * 8));
                    HIP_TRY(r, hipMalloc(&r->pidx[k], cap * 4));
                }
                r->pcap = cap;
            }
            {
                Prof pr(r, "particles_keys");
                launch_particles(r->stream, f, r->view, r->offsets, r->bs, 1, seed, r->depth, S, r->lut, r->stats,
                                 grid, r->pkeys[0], r->pidx[0], (int)g_tuning.particles_per_thread, nullptr,
                                 (int)g_tuning.particle_block);
            }
            int res = 0;
            // keys are written at their particle index, so the low word is already in order
            if (onesweep_sort(r->os, r->stream, r->pkeys, r->pidx, total, res, r->err,
                              r->profiling ? &r->pending : nullptr, 32) != 0)
                return -1;
            r->psort_result = res;
            {
                Prof pr(r, "first_hit");
                launch_first_hit(r->stream, r->pidx[res], total, r->first, S);
                launch_build_depth(r->stream, r->pkeys[res], r->first, S, r->offsets, (uint32_t)r->n, r->depth);
            }
        }
        r->last_np = total;
        {
            Prof pr(r, "resolve");
            launch_resolve(r->stream, f, r->depth, S, r->bs.skip, 1, r->colour, r->accum, keep ? 0 : 1, r->stats,
                           r->mom, r->zmin);
        }
        if (keep) r->depth_dirty = true;
        r->have_last = keep;
    }
    HIP_TRY(r, hipGetLastError());
    r->accumulated += count;
    return 0;
}

}  // namespace

// ====================================================================== C ABI
void renderer_device_buffers(GpsRenderer *r, const float4 **accum, const float4 **mom, const float **zmin) {
    *accum = r->accum;
    *mom = r->mom;
    *zmin = r->zmin;
}

extern "C" {

void gps_render_params_default(GpsRenderParams *o) {
    std::memset(o, 0, sizeof *o);
    o->spp_side = 2;
    o->density_scale = 1.0f;
    o->max_points_per_splat = 1u << 16;
    o->near = 0.05f;
    o->background[0] = 0.05f;
    o->background[1] = 0.05f;
    o->background[2] = 0.08f;
    o->exposure = 1.0f;
    o->srgb_output = 1;
    o->exact_radius = 0;
    o->sh_degree = 3;
    o->path = GPS_PATH_PARTICLE;
    o->method = GPS_METHOD_GPS;
    o->pbvr.calibration = GPS_CAL_PER_SPLAT_FOOTPRINT;
    o->pbvr.radial_correction = 0;
    o->pbvr.centre_depth = 0;
    o->pbvr.low_pass_jitter = 1;
    o->pbvr.base_k = 512.0f;
    o->pbvr.reference_pixel_length = 0.01f;
}

const char *gps_renderer_last_error(const GpsRenderer *r) { return r ? r->err.c_str() : g_create_error.c_str(); }

int gps_renderer_create(uint32_t width, uint32_t height, uint32_t spp_side, int device, GpsRenderer **out) {
    *out = nullptr;
    if (width == 0 || height == 0) {
        g_create_error = "width/height must be > 0";
        return -1;
    }
    GpsRenderer *r = new GpsRenderer();
    r->device = device;
    r->W = width;
    r->H = height;
    r->spp_side = spp_side < 1 ? 1 : (spp_side > 4 ? 4 : spp_side);
    auto bail = [&](const std::string &m) {
        g_create_error = m;
        gps_renderer_destroy(r);
        return -1;
    };
    if (hipSetDevice(device) != hipSuccess) return bail("hipSetDevice failed (no GPU?)");
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return bail("hipStreamCreate failed");
    std::vector<float> lut(LUT_O * LUT_V);
    make_lut(lut.data());
    if (hipMalloc(&r->lut, lut.size() * sizeof(float)) != hipSuccess) return bail("hipMalloc(lut) failed");
    if (hipMemcpy(r->lut, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bail("lut upload failed");
    if (hipMalloc(&r->stats, ST_WORDS * sizeof(uint64_t)) != hipSuccess) return bail("hipMalloc(stats) failed");
    if (hipHostMalloc((void **)&r->stats_staging, ST_WORDS * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
        return bail("hipHostMalloc(stats staging) failed");
    if (hipEventCreateWithFlags(&r->stats_event, hipEventDisableTiming) != hipSuccess) return bail("hipEventCreate");
    if (hipMemsetAsync(r->stats, 0, ST_WORDS * sizeof(uint64_t), r->stream) != hipSuccess) return bail("stats clear");
    if (hipMalloc(&r->bs.tickets, MAX_BATCH * sizeof(uint32_t)) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc(&r->bs.totals, MAX_BATCH * sizeof(uint32_t)) != hipSuccess) return bail("hipMalloc failed");
    if (hipMalloc(&r->bs.skip, MAX_BATCH * sizeof(uint32_t)) != hipSuccess) return bail("hipMalloc failed");
    if (hipMemsetAsync(r->bs.tickets, 0, MAX_BATCH * sizeof(uint32_t), r->stream) != hipSuccess) return bail("memset");
    if (hipMemsetAsync(r->bs.totals, 0, MAX_BATCH * sizeof(uint32_t), r->stream) != hipSuccess) return bail("memset");
    if (hipMemsetAsync(r->bs.skip, 0, MAX_BATCH * sizeof(uint32_t), r->stream) != hipSuccess) return bail("memset");
    if (alloc_frame(r) != 0) return bail(r->err);
    if (gps_renderer_set_gaussians(r, nullptr, 0, nullptr, 0) != 0) return bail(r->err);
    r->sorted = sorted_create();
    r->os = onesweep_create();
    if (hipStreamSynchronize(r->stream) != hipSuccess) return bail("stream sync failed");
    *out = r;
    return 0;
}

void gps_renderer_destroy(GpsRenderer *r) {
    if (!r) return;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    drain_profile(r);
    if (r->sorted) sorted_destroy(r->sorted);
    if (r->os) onesweep_destroy(r->os);
    for (int k = 0; k < 2; k++) {
        dfree(r->pkeys[k]);
        dfree(r->pidx[k]);
    }
    dfree(r->first);
    dfree(r->splats);
    dfree(r->sh);
    dfree(r->lut);
    dfree(r->view);
    dfree(r->perm);
    dfree(r->cinfo);
    dfree(r->colour);
    dfree(r->offsets);
    dfree(r->depth);
    dfree(r->bs.tickets);
    dfree(r->bs.status);
    dfree(r->bs.totals);
    dfree(r->bs.skip);
    dfree(r->accum);
    dfree(r->mom);
    dfree(r->zmin);
    dfree(r->stats);
    if (r->stats_staging) (void)hipHostFree(r->stats_staging);
    if (r->stats_event) (void)hipEventDestroy(r->stats_event);
    dfree(r->composite_buf);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

int gps_renderer_set_gaussians(GpsRenderer *r, const GpsGaussian *g, uint64_t n, const float *sh_rest,
                               uint32_t degree) {
    if (n > 0 && !g) return fail(r, "null Gaussian pointer");
    if (n >= MAX_SPLATS) return fail(r, "too many Gaussians for one dispatch");
    uint32_t deg = 0, stride = 0;
    if (sh_rest && degree >= 1 && n > 0) {
        deg = degree > 3 ? 3 : degree;
        stride = (deg + 1) * (deg + 1) - 1;
    }
    HIP_TRY(r, hipSetDevice(r->device));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    // allocate the new scene first so that a failure leaves the old one in place (gps.rs:682)
    float4 *splats = nullptr, *view = nullptr, *colour = nullptr;
    float *sh = nullptr;
    uint2 *cinfo = nullptr;
    uint32_t *perm = nullptr;
    uint64_t nn = n ? n : 1;
    auto cleanup = [&]() {
        dfree(splats);
        dfree(view);
        dfree(colour);
        dfree(sh);
        dfree(cinfo);
        dfree(perm);
    };
    if (hipMalloc(&splats, nn * 64) != hipSuccess || hipMalloc(&view, nn * VIEW_VEC4 * sizeof(float4)) != hipSuccess ||
        hipMalloc(&colour, nn * sizeof(float4)) != hipSuccess || hipMalloc(&cinfo, nn * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&perm, (nn + 3) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&sh, (stride ? nn * 3 * stride : 16) * sizeof(float)) != hipSuccess) {
        cleanup();
        return fail(r, "out of device memory for the scene");
    }
    if (n) {
        if (hipMemcpyAsync(splats, g, n * 64, hipMemcpyHostToDevice, r->stream) != hipSuccess ||
            (stride && hipMemcpyAsync(sh, sh_rest, n * 3 * stride * sizeof(float), hipMemcpyHostToDevice,
                                      r->stream) != hipSuccess)) {
            cleanup();
            return fail(r, "scene upload failed");
        }
    }
    dfree(r->splats);
    dfree(r->view);
    dfree(r->colour);
    dfree(r->sh);
    dfree(r->cinfo);
    dfree(r->perm);
    r->splats = splats;
    r->view = view;
    r->perm = perm;
    r->colour = colour;
    r->sh = sh;
    r->cinfo = cinfo;
    r->n = n;
    r->sh_max_degree = deg;
    r->sh_stride = stride;
    double c[3] = {0, 0, 0};
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) c[k] = c[k] + (double)g[i].pos_opacity[k];
    for (int k = 0; k < 3; k++) r->centroid[k] = n ? c[k] / (double)n : 0.0;
    if (alloc_batch(r) != 0) return -1;
    if (gps_renderer_reset_accum(r) != 0) return -1;
    HIP_TRY(r, hipStreamSynchronize(r->stream));   // the caller may free the host arrays after return
    return 0;
}

int gps_renderer_resize(GpsRenderer *r, uint32_t width, uint32_t height, uint32_t spp_side) {
    spp_side = spp_side < 1 ? 1 : (spp_side > 4 ? 4 : spp_side);
    if (width == r->W && height == r->H && spp_side == r->spp_side) return 0;
    if (width == 0 || height == 0) return fail(r, "width/height must be > 0");
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    r->W = width;
    r->H = height;
    r->spp_side = spp_side;
    if (alloc_frame(r) != 0) return -1;
    if (alloc_batch(r) != 0) return -1;
    return gps_renderer_reset_accum(r);
}

int gps_renderer_reset_accum(GpsRenderer *r) {
    launch_accum_clear(r->stream, r->accum, (uint64_t)r->W * r->H);
    launch_features_clear(r->stream, r->mom, r->zmin, (uint64_t)r->W * r->H);
    launch_stats_clear(r->stream, r->stats, ST_SKIPPED, 1);
    r->accumulated = 0;
    HIP_TRY(r, hipGetLastError());
    return 0;
}

int gps_renderer_reset_stats(GpsRenderer *r) {
    launch_stats_clear(r->stream, r->stats, 0, ST_WORDS);
    HIP_TRY(r, hipGetLastError());
    return 0;
}

static int enqueue_ensembles(GpsRenderer *r, hipStream_t st, const GpsCamera *cam, const GpsRenderParams *p,
                             uint32_t first_seed, uint32_t count);

int gps_renderer_render_ensembles(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p,
                                  uint32_t first_seed, uint32_t count) {
    if (count == 0) return 0;
    if (check_params(r, p) != 0) return -1;
    HIP_TRY(r, hipSetDevice(r->device));
    return enqueue_ensembles(r, r->stream, cam, p, first_seed, count);
}

// Enqueues the ensembles on `st` (the renderer's stream, or a batch stream under capture).
static int enqueue_ensembles(GpsRenderer *r, hipStream_t st, const GpsCamera *cam, const GpsRenderParams *p,
                      uint32_t first_seed, uint32_t count) {
    r->enqueue_stream = st;
    GpsFrame f;
    pack_frame(r, cam, p, f);
    const uint64_t S = sub_count(r);
    const bool use_particles = (p->path != GPS_PATH_PRIMITIVE || p->method != GPS_METHOD_GPS) && r->n > 0;
    const bool keep = (r->debug & GPS_DEBUG_KEEP_BUFFERS) != 0;
    const bool sorted = use_particles && p->path == GPS_PATH_PARTICLE_SORTED;
    // count-only splat order: the packed-minimum path only (the sorted path keys particles by their index in
    // splat order), and not when the offsets are kept for inspection
    uint32_t *perm = ((f.flags & FLAG_FAST) && use_particles && !sorted && !keep) ? r->perm : nullptr;
    if (r->n > 0) {
        Prof pr(r, "view");
        if (perm) launch_zero_u32(st, perm + perm_counter((uint32_t)r->n), 2);
        launch_view(st, f, r->splats, r->sh, r->view, r->cinfo, r->colour, perm);
    }
    if (sorted) {
        if (st != r->stream) return fail(r, "the sorted particle path cannot be recorded into a batch graph");
        return render_sorted_particles(r, f, first_seed, count);
    }
    const int grid = grid_for_particles(r);
    uint32_t done = 0;
    const uint32_t cap = ensembles_per_launch(r);
    while (done < count) {
        uint32_t B = count - done;
        if (B > cap) B = cap;
        if (keep) B = 1;
        const uint32_t seed0 = first_seed + done;
        if (r->depth_dirty) {
            Prof pr(r, "clear");
            launch_fill_u64(st, r->depth, (uint64_t)B * S, EMPTY64);
            r->depth_dirty = false;
        }
        if (use_particles) {
            {
                Prof pr(r, "count_scan");
                launch_count_scan(st, f, r->cinfo, seed0, B, r->offsets, r->bs, r->tiles, r->epoch, r->stats, perm);
            }
            r->epoch = (r->epoch + 1u) & ((1u << 26) - 1u);
            if (r->epoch == 0) r->epoch = 1;
            {
                Prof pr(r, "particles");
                launch_particles(st, f, r->view, r->offsets, r->bs, B, seed0, r->depth, S, r->lut, r->stats,
                                 grid, nullptr, nullptr, (int)g_tuning.particles_per_thread, perm,
                                 (int)g_tuning.particle_block);
            }
        } else if (r->n > 0) {
            Prof pr(r, "splat_pass");
            for (uint32_t b = 0; b < B; b++)
                launch_splat_pass(st, f, r->view, r->cinfo, seed0 + b, r->depth + (uint64_t)b * S, r->lut,
                                  r->stats);
        }
        {
            Prof pr(r, "resolve");
            launch_resolve(st, f, r->depth, S, use_particles ? r->bs.skip : nullptr, B, r->colour, r->accum,
                           keep ? 0 : 1, r->stats, r->mom, r->zmin);
        }
        if (keep) r->depth_dirty = true;
        r->have_last = keep && use_particles;
        done += B;
    }
    HIP_TRY(r, hipGetLastError());
    r->accumulated += count;
    return 0;
}

int gps_renderer_render_sorted(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p, uint32_t as_ensembles) {
    if (r->n == 0) return fail(r, "empty scene");
    if (check_params(r, p) != 0) return -1;
    GpsRenderParams q = *p;
    q.method = GPS_METHOD_GPS;
    GpsFrame f;
    pack_frame(r, cam, &q, f);
    if (sorted_render(r->sorted, r->stream, f, r->splats, r->sh, r->accum, r->stats, 0, r->err,
                      r->profiling ? &r->pending : nullptr) != 0)
        return -1;
    r->accumulated = as_ensembles > 1 ? as_ensembles : 1;
    return 0;
}

int gps_renderer_probe_demand(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p) {
    if (r->n == 0) return fail(r, "empty scene");
    GpsFrame f;
    pack_frame(r, cam, p, f);
    f.method = GPS_METHOD_GPS;
    return sorted_render(r->sorted, r->stream, f, r->splats, r->sh, r->accum, r->stats, 1, r->err,
                         r->profiling ? &r->pending : nullptr);
}

int gps_renderer_read_sorted(GpsRenderer *r, uint32_t *keys, uint32_t *vals) {
    return sorted_read(r->sorted, r->stream, (uint32_t)r->n, keys, vals, r->err);
}

int gps_renderer_composite(GpsRenderer *r, const GpsRenderParams *p, uint32_t out_w, uint32_t out_h, uint8_t *rgba8) {
    if (out_w == 0 || out_h == 0) return fail(r, "empty composite target");
    uint64_t px = (uint64_t)out_w * out_h;
    if (px > r->composite_cap) {
        dfree(r->composite_buf);
        HIP_TRY(r, hipMalloc(&r->composite_buf, px * 4));
        r->composite_cap = px;
    }
    {
        Prof pr(r, "composite");
        launch_composite(r->stream, r->accum, r->W, r->H, p->exposure, p->srgb_output ? 1u : 0u, out_w, out_h,
                         r->composite_buf);
    }
    HIP_TRY(r, hipMemcpyAsync(rgba8, r->composite_buf, px * 4, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    return 0;
}

int gps_renderer_read_accum(GpsRenderer *r, float *out) {
    HIP_TRY(r, hipMemcpyAsync(out, r->accum, (uint64_t)r->W * r->H * sizeof(float4), hipMemcpyDeviceToHost,
                              r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    return 0;
}

int gps_renderer_set_features(GpsRenderer *r, int enable) {
    HIP_TRY(r, hipSetDevice(r->device));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    const uint64_t X = (uint64_t)r->W * r->H;
    if (!enable) {
        dfree(r->mom);
        dfree(r->zmin);
        r->features = false;
        return 0;
    }
    if (!r->features) {
        r->features = true;
        HIP_TRY(r, hipMalloc(&r->mom, X * sizeof(float4)));
        HIP_TRY(r, hipMalloc(&r->zmin, X * sizeof(float)));
    }
    launch_features_clear(r->stream, r->mom, r->zmin, X);
    HIP_TRY(r, hipGetLastError());
    return 0;
}

int gps_renderer_set_cull(GpsRenderer *r, int enable) {
    r->cull = enable != 0;
    return 0;
}

int gps_renderer_read_features(GpsRenderer *r, float *moments, float *zmin) {
    if (!r->features) return fail(r, "features are not enabled");
    const uint64_t X = (uint64_t)r->W * r->H;
    if (moments) HIP_TRY(r, hipMemcpyAsync(moments, r->mom, X * sizeof(float4), hipMemcpyDeviceToHost, r->stream));
    if (zmin) HIP_TRY(r, hipMemcpyAsync(zmin, r->zmin, X * sizeof(float), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    return 0;
}

static void stats_from_words(const uint64_t *w, GpsStats *out) {
    std::memset(out, 0, sizeof *out);
    out->points = w[ST_POINTS];
    out->truncated_splats = w[ST_TRUNC];
    out->dropped_points = w[ST_DROPPED];
    out->visible_splats = w[ST_VISIBLE];
    out->orphan_subpixels = w[ST_ORPHAN];
    out->candidates = w[ST_CAND];
    out->undercovered_splats = w[ST_UNDER];
    out->skipped_ensembles = w[ST_SKIPPED];
    out->ensembles = w[ST_ENSEMBLES];
    out->probe_demand = w[ST_PROBE];
}

int gps_renderer_read_stats(GpsRenderer *r, GpsStats *out) {
    uint64_t w[ST_WORDS];
    HIP_TRY(r, hipMemcpyAsync(w, r->stats, sizeof w, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    stats_from_words(w, out);
    return 0;
}

// Delivers a finished non-blocking readback into stats_latest (no wait).
static void poll_stats(GpsRenderer *r) {
    if (!r->stats_pending || hipEventQuery(r->stats_event) != hipSuccess) return;
    stats_from_words(r->stats_staging, &r->stats_latest);
    r->stats_latest.issue = r->stats_pending_issue;
    r->stats_pending = false;
}

int gps_renderer_request_stats(GpsRenderer *r) {
    poll_stats(r);
    if (r->stats_pending) return 0;   // one readback in flight at a time
    HIP_TRY(r, hipMemcpyAsync(r->stats_staging, r->stats, ST_WORDS * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              r->stream));
    HIP_TRY(r, hipEventRecord(r->stats_event, r->stream));
    r->stats_pending = true;
    r->stats_pending_issue = ++r->stats_issued;
    return 0;
}

int gps_renderer_latest_stats(GpsRenderer *r, GpsStats *out) {
    poll_stats(r);
    *out = r->stats_latest;
    return 0;
}

uint32_t gps_renderer_stats_issued(const GpsRenderer *r) { return r->stats_issued; }

int gps_renderer_wait_idle(GpsRenderer *r) {
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    return 0;
}

uint32_t gps_renderer_submitted(const GpsRenderer *r) { return r->accumulated; }

uint32_t gps_renderer_accumulated(GpsRenderer *r) {
    GpsStats s;
    if (gps_renderer_read_stats(r, &s) != 0) return r->accumulated;
    uint64_t sk = s.skipped_ensembles;
    return sk >= r->accumulated ? 0u : r->accumulated - (uint32_t)sk;
}

void gps_renderer_info(const GpsRenderer *r, uint32_t *w, uint32_t *h, uint32_t *spp, uint64_t *n, uint32_t *deg) {
    if (w) *w = r->W;
    if (h) *h = r->H;
    if (spp) *spp = r->spp_side;
    if (n) *n = r->n;
    if (deg) *deg = r->sh_max_degree;
}

void gps_estimate_memory(uint32_t width, uint32_t height, uint32_t spp_side, uint64_t n_splats, uint64_t *dw,
                         uint64_t *ac, uint64_t *sp) {
    uint64_t sub = (uint64_t)width * height * (uint64_t)(spp_side * spp_side);
    if (dw) *dw = 8 * sub;
    if (ac) *ac = 16ull * width * height;
    if (sp) *sp = 164ull * n_splats;
}

void gps_radius_lut(float *out) { make_lut(out); }

int gps_renderer_set_debug(GpsRenderer *r, uint32_t flags) {
    r->debug = flags;
    return 0;
}

int gps_renderer_read_counts(GpsRenderer *r, uint32_t *counts, uint32_t *offsets, uint32_t *total, uint32_t *skipped) {
    if (!r->have_last) return fail(r, "no kept ensemble: enable GPS_DEBUG_KEEP_BUFFERS and render the particle path");
    uint64_t n = r->n;
    std::vector<uint32_t> off(n ? n : 1);
    uint32_t t = 0, sk = 0;
    HIP_TRY(r, hipMemcpyAsync(off.data(), r->offsets, n * 4, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipMemcpyAsync(&t, r->bs.totals, 4, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipMemcpyAsync(&sk, r->bs.skip, 4, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    if (offsets) std::memcpy(offsets, off.data(), n * 4);
    if (counts && !sk)
        for (uint64_t i = 0; i < n; i++) counts[i] = (i + 1 < n ? off[i + 1] : t) - off[i];
    if (total) *total = t;
    if (skipped) *skipped = sk;
    return 0;
}

int gps_renderer_read_depth_id(GpsRenderer *r, uint64_t *out) {
    if (!(r->debug & GPS_DEBUG_KEEP_BUFFERS)) return fail(r, "enable GPS_DEBUG_KEEP_BUFFERS first");
    HIP_TRY(r, hipMemcpyAsync(out, r->depth, sub_count(r) * 8, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    return 0;
}

int gps_renderer_read_particle_counters(GpsRenderer *r, uint64_t *out) {
    uint64_t w[ST_WORDS];
    HIP_TRY(r, hipMemcpyAsync(w, r->stats, sizeof w, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    out[0] = w[ST_HITS];
    out[1] = w[ST_QPEAK];
    out[2] = PARTICLE_WAVE_CAP;
    out[3] = w[ST_WPEAK];
    out[4] = PARTICLE_WAVE_CAP;
    return 0;
}

int gps_renderer_read_particle_keys(GpsRenderer *r, uint64_t *keys, uint32_t *subpixels, uint64_t max,
                                    uint64_t *count) {
    if (count) *count = r->last_np;
    uint64_t n = r->last_np < max ? r->last_np : max;
    if (n == 0) return 0;
    HIP_TRY(r, hipMemcpyAsync(keys, r->pkeys[r->psort_result], n * 8, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipMemcpyAsync(subpixels, r->pidx[r->psort_result], n * 4, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    return 0;
}

int gps_renderer_profile(GpsRenderer *r, int enable) {
    HIP_TRY(r, hipStreamSynchronize(r->stream));
    drain_profile(r);
    r->profiling = enable != 0;
    if (enable) r->times.clear();
    return 0;
}

int gps_renderer_profile_read(GpsRenderer *r, GpsKernelTime *out, int max) {
    if (hipStreamSynchronize(r->stream) != hipSuccess) return -1;
    drain_profile(r);
    int i = 0;
    for (auto &t : r->times) {
        if (i < max && out) {
            std::memset(&out[i], 0, sizeof out[i]);
            std::snprintf(out[i].name, sizeof out[i].name, "%s", t.first.c_str());
            out[i].launches = t.second.first;
            out[i].total_ms = t.second.second;
        }
        i++;
    }
    return i;
}

// -------------------------------------------------------------- batches of views (one hipGraph per batch)
// Items of different renderers are independent: the batch forks into up to BATCH_LANES streams (graph
// branches under capture), one per renderer in order of first appearance (round robin), and joins them.
constexpr int BATCH_LANES = 4;

// A fused item's buffers, owned by the batch and kept for its next render when large enough.
struct FusedBuf {
    float4 *view = nullptr, *colour = nullptr, *mom = nullptr;
    uint2 *cinfo = nullptr;
    uint32_t *perm = nullptr, *offsets = nullptr;
    uint32_t *small = nullptr;   // tickets, totals, skip: MAX_BATCH words each, whatever a batch's round size
    uint64_t *depth = nullptr, *status = nullptr;
    float *zmin = nullptr;
    uint64_t n_cap = 0, off_cap = 0, depth_cap = 0, status_cap = 0, ens_cap = 0, x_cap = 0;   // in elements
    bool depth_clean = false;   // every depth|id word is all ones (a finished round leaves it so)
};

struct GpsBatch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t lanes[BATCH_LANES] = {};
    hipEvent_t fork = nullptr, join[BATCH_LANES] = {};
    std::vector<float4 *> outs;
    std::vector<uint64_t> out_px;
    uint64_t *item_stats = nullptr;   // ST_WORDS counters per item
    uint32_t item_stats_cap = 0;
    std::vector<uint32_t> item_w, item_h;   // sizes of the last render's items
    uint32_t *composite_buf = nullptr;
    uint64_t composite_cap = 0;
    double gpu_ms = 0.0, record_ms = 0.0;
    uint64_t nodes = 0;
    // fused path: per-item buffers, the device item table and the block tables
    std::vector<FusedBuf> fbuf;
    FusedItem *fitems = nullptr;
    uint32_t fitems_cap = 0;
    uint32_t *tables = nullptr;
    uint64_t tables_cap = 0;
    uint32_t epoch = 1;
};

int gps_batch_create(int device, GpsBatch **out) {
    *out = nullptr;
    GpsBatch *b = new GpsBatch();
    b->device = device;
    bool ok = hipSetDevice(device) == hipSuccess &&
              hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int l = 0; ok && l < BATCH_LANES; l++)
        ok = hipStreamCreateWithFlags(&b->lanes[l], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        g_create_error = "batch stream creation failed";
        gps_batch_destroy(b);
        return -1;
    }
    *out = b;
    return 0;
}

void gps_batch_destroy(GpsBatch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (int l = 0; l < BATCH_LANES; l++) {
        if (b->lanes[l]) {
            (void)hipStreamSynchronize(b->lanes[l]);
            (void)hipStreamDestroy(b->lanes[l]);
        }
        if (b->join[l]) (void)hipEventDestroy(b->join[l]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto *o : b->outs)
        if (o) (void)hipFree(o);
    for (FusedBuf &f : b->fbuf) {
        for (void *q : {(void *)f.view, (void *)f.colour, (void *)f.mom, (void *)f.cinfo, (void *)f.perm,
                        (void *)f.offsets, (void *)f.small, (void *)f.depth, (void *)f.status, (void *)f.zmin})
            if (q) (void)hipFree(q);
    }
    if (b->fitems) (void)hipFree(b->fitems);
    if (b->tables) (void)hipFree(b->tables);
    if (b->item_stats) (void)hipFree(b->item_stats);
    if (b->composite_buf) (void)hipFree(b->composite_buf);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}


extern "C++" {
namespace {

// (re)allocates p for `need` elements when its capacity is smaller; false when out of memory
template <typename T>
bool grow(T *&p, uint64_t &cap, uint64_t need) {
    if (need <= cap && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, (need ? need : 1) * sizeof(T)) != hipSuccess) return false;
    cap = need;
    return true;
}

// Emitter class of an item on the particle path, or -1 when it cannot be fused (the splat-parallel GPS path).
int fused_mk(const GpsRenderParams *p) {
    if (p->method == GPS_METHOD_GPS) return p->path == GPS_PATH_PARTICLE ? 0 : -1;
    return p->method == GPS_METHOD_SCREEN_OCCUPANCY ? 2 : 1;
}

}  // namespace
}  // extern "C++"

// Fused batch: every item renders through the same few launches.  Items get buffers of their own (view cache,
// offsets, depth|id buffers for `round_ens` ensembles, image, counters), so items that share a renderer are
// independent.  Launches: clear, view (all items), then per round of up to round_ens ensembles of every item: count
// and scan (every (item, ensemble) tile), per emitter class an assignment of the particle grid to items and the
// particle kernel, and resolve (every item's pixels).  The particle grid is split among the items in proportion
// to their chunks, each item's blocks running the single-view kernel's body over its own chunks.
static int fused_batch_render(GpsBatch *b, const GpsBatchItem *items, uint32_t n, bool use_graph) {
    const int ppt = (int)g_tuning.particles_per_thread, block = (int)g_tuning.particle_block;
    const uint32_t chunk = (uint32_t)(ppt * block);
    // rounds: ensembles per round from the tuning, else as many as keep the largest item's depth|id buffers
    // within the depth window
    uint64_t max_s = 1, max_x = 1;
    uint32_t max_count = 0;
    for (uint32_t i = 0; i < n; i++) {
        const GpsRenderer *r = items[i].renderer;
        max_s = std::max<uint64_t>(max_s, sub_count(r));
        max_x = std::max<uint64_t>(max_x, (uint64_t)r->W * r->H);
        max_count = std::max(max_count, items[i].count);
    }
    uint32_t round_ens = g_tuning.batch_ensembles;
    if (!round_ens) {
        const uint64_t w = ((uint64_t)g_tuning.depth_window_mib << 20) / (8 * max_s);
        round_ens = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(w, 1), MAX_BATCH);
    }
    const uint32_t rounds = (max_count + round_ens - 1) / round_ens;
    // buffers and descriptors
    if (b->fbuf.size() < n) b->fbuf.resize(n);
    std::vector<FusedItem> fi(n);
    bool mk_used[3] = {false, false, false};
    for (uint32_t i = 0; i < n; i++) {
        const GpsBatchItem &it = items[i];
        GpsRenderer *r = it.renderer;
        FusedBuf &q = b->fbuf[i];
        FusedItem &d = fi[i];
        pack_frame(r, it.camera, it.params, d.f);
        const uint64_t N = r->n, S = sub_count(r), X = (uint64_t)r->W * r->H;
        const uint32_t tiles = (uint32_t)((N + SCAN_TILE - 1) / SCAN_TILE);
        uint64_t cap_n = q.n_cap;   // the view cache, colours, count info and splat order share one capacity
        bool ok = true;
        if (N > q.n_cap) {
            uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
            ok = grow(q.view, c0, (uint64_t)VIEW_VEC4 * N) && grow(q.colour, c1, N) && grow(q.cinfo, c2, N) && grow(q.perm, c3, N + 3);
            cap_n = ok ? N : 0;
        }
        q.n_cap = cap_n;
        const uint64_t ecap = round_ens;
        ok = ok && grow(q.offsets, q.off_cap, ecap * N);
        const bool depth_new = q.depth_cap < ecap * S || !q.depth;
        ok = ok && grow(q.depth, q.depth_cap, ecap * S);
        const bool status_new = q.status_cap < ecap * tiles || !q.status;
        ok = ok && grow(q.status, q.status_cap, ecap * std::max<uint32_t>(tiles, 1));
        const bool small_new = !q.small;
        ok = ok && grow(q.small, q.ens_cap, 3 * (uint64_t)MAX_BATCH);
        if (ok && r->features && (X > q.x_cap || !q.mom || !q.zmin)) {
            uint64_t c0 = 0, c1 = 0;
            ok = grow(q.mom, c0, X) && grow(q.zmin, c1, X);
            q.x_cap = ok ? X : 0;
        }
        if (!ok) {
            g_create_error = "out of device memory for a fused batch";
            return -1;
        }
        if (depth_new) q.depth_clean = false;
        if ((status_new && hipMemsetAsync(q.status, 0, q.status_cap * sizeof(uint64_t), b->stream) != hipSuccess) ||
            (small_new && hipMemsetAsync(q.small, 0, q.ens_cap * sizeof(uint32_t), b->stream) != hipSuccess)) {
            g_create_error = "fused batch buffer clear failed";
            return -1;
        }
        if (!q.depth_clean) {
            launch_fill_u64(b->stream, q.depth, q.depth_cap, EMPTY64);
            q.depth_clean = true;
        }
        d.splats = r->splats;
        d.sh = r->sh;
        d.lut = r->lut;
        d.view = q.view;
        d.cinfo = q.cinfo;
        d.colour = q.colour;
        d.perm = (d.f.flags & FLAG_FAST) ? q.perm : nullptr;
        d.offsets = q.offsets;
        d.depth = q.depth;
        d.bs.tickets = q.small;   // zero between scans: the last tile of an ensemble resets its word
        d.bs.totals = q.small + MAX_BATCH;
        d.bs.skip = q.small + 2 * MAX_BATCH;
        d.bs.status = q.status;
        d.accum = b->outs[i];
        d.mom = r->features ? q.mom : nullptr;
        d.zmin = r->features ? q.zmin : nullptr;
        d.stats = b->item_stats + (size_t)i * ST_WORDS;
        d.sub_count = S;
        d.tiles = tiles;
        d.first_seed = it.first_seed;
        d.mk = (uint32_t)fused_mk(it.params);
        mk_used[d.mk] = true;
    }
    // block tables: view (per item ceil(N / 256)), resolve (ceil(X / 256)), count-scan tiles per round, and the
    // particle grid's assignment (written on the device)
    GpsRenderer *r0 = items[0].renderer;
    const uint32_t grid = std::max<uint32_t>((uint32_t)grid_for_particles(r0), n);
    std::vector<uint32_t> tab;
    auto blocks_of = [&](auto per_item, std::vector<uint32_t> &blk0) {
        std::vector<uint32_t> bi;
        for (uint32_t i = 0; i < n; i++) {
            blk0.push_back((uint32_t)bi.size());
            const uint64_t nb = per_item(i);
            for (uint64_t k = 0; k < nb; k++) bi.push_back(i);
        }
        return bi;
    };
    std::vector<uint32_t> v0, x0;
    std::vector<uint32_t> vb = blocks_of([&](uint32_t i) { return (items[i].renderer->n + 255) / 256; }, v0);
    std::vector<uint32_t> xb = blocks_of([&](uint32_t i) {
        return ((uint64_t)items[i].renderer->W * items[i].renderer->H + 255) / 256; }, x0);
    // per round: each item's ensembles (0 once it has done its count) and the scan's (item, ensemble) tiles
    std::vector<uint32_t> ne_all((size_t)rounds * n);
    std::vector<std::vector<uint32_t>> cs(rounds);
    for (uint32_t rd = 0; rd < rounds; rd++)
        for (uint32_t i = 0; i < n; i++) {
            const uint64_t e0 = (uint64_t)rd * round_ens;
            const uint32_t ne = items[i].count > e0 ? (uint32_t)std::min<uint64_t>(round_ens, items[i].count - e0) : 0;
            ne_all[(size_t)rd * n + i] = ne;
            for (uint32_t e = 0; e < ne; e++)
                for (uint32_t t = 0; t < fi[i].tiles; t++) cs[rd].push_back((i << 8) | e);
        }
    // layout: vb | v0 | xb | x0 | ne_all | cs[0] | cs[1] ... | blk_item (grid) | blk_rank (grid) | item_nblk (n)
    auto put = [&](const std::vector<uint32_t> &v) {
        const uint64_t at = tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        return at;
    };
    const uint64_t o_vb = put(vb), o_v0 = put(v0), o_xb = put(xb), o_x0 = put(x0), o_ne = put(ne_all);
    std::vector<uint64_t> o_cs(rounds);
    for (uint32_t rd = 0; rd < rounds; rd++) o_cs[rd] = put(cs[rd]);
    const uint64_t o_bi = tab.size(), o_br = o_bi + grid, o_nb = o_br + grid;
    tab.resize(o_nb + n, 0u);
    if (!grow(b->tables, b->tables_cap, tab.size())) {
        g_create_error = "out of device memory for a fused batch";
        return -1;
    }
    if (b->fitems_cap < n) {
        if (b->fitems) (void)hipFree(b->fitems);
        b->fitems = nullptr;
        b->fitems_cap = 0;
        if (hipMalloc(&b->fitems, (size_t)n * sizeof(FusedItem)) != hipSuccess) {
            g_create_error = "out of device memory for a fused batch";
            return -1;
        }
        b->fitems_cap = n;
    }
    if (hipMemcpyAsync(b->tables, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->fitems, fi.data(), (size_t)n * sizeof(FusedItem), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        g_create_error = "fused batch upload failed";
        return -1;
    }
    uint32_t *T = b->tables;
    auto t0 = std::chrono::steady_clock::now();
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph && hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        g_create_error = "hipStreamBeginCapture failed";
        return -1;
    }
    for (uint32_t i = 0; i < n; i++) b->fbuf[i].depth_clean = false;   // until the batch has finished
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    if (!use_graph) (void)hipEventRecord(e0, b->stream);
    hipStream_t st = b->stream;
    launch_fused_clear(st, b->fitems, n, (uint32_t)max_x);
    launch_fused_view(st, b->fitems, T + o_vb, T + o_v0, (uint32_t)vb.size());
    for (uint32_t rd = 0; rd < rounds; rd++) {
        const uint32_t *ne = T + o_ne + (size_t)rd * n;
        const uint32_t ens0 = rd * round_ens;   // seeds are first_seed + ensemble modulo 2^32
        launch_fused_count_scan(st, b->fitems, T + o_cs[rd], (uint32_t)cs[rd].size(), ens0, b->epoch);
        b->epoch = (b->epoch + 1u) & ((1u << 26) - 1u);
        if (b->epoch == 0) b->epoch = 1;
        for (uint32_t mk = 0; mk < 3; mk++) {
            if (!mk_used[mk]) continue;
            launch_fused_assign(st, b->fitems, n, ne, mk, chunk, grid, T + o_bi, T + o_br, T + o_nb);
            launch_fused_particles(st, b->fitems, mk, ne, ens0, T + o_bi, T + o_br, T + o_nb, (int)grid, ppt, block);
        }
        launch_fused_resolve(st, b->fitems, ne, T + o_xb, T + o_x0, (uint32_t)xb.size());
    }
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (!use_graph) (void)hipEventRecord(e1, b->stream);
    b->nodes = 0;
    if (use_graph) {
        if (hipStreamEndCapture(b->stream, &graph) != hipSuccess) rc = -1;
        if (rc == 0 && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = -1;
        size_t nn = 0;
        if (rc == 0 && hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess) b->nodes = nn;
    }
    b->record_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == 0 && use_graph) {
        (void)hipEventRecord(e0, b->stream);
        if (hipGraphLaunch(exec, b->stream) != hipSuccess) rc = -1;
        (void)hipEventRecord(e1, b->stream);
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) rc = -1;
    if (e0) {
        float ms = 0.0f;
        if (rc == 0) (void)hipEventElapsedTime(&ms, e0, e1);
        b->gpu_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != 0) {
        if (g_create_error.empty()) g_create_error = "fused batch failed";
        return -1;
    }
    for (uint32_t i = 0; i < n; i++) b->fbuf[i].depth_clean = true;   // every round's resolve reset them
    // each renderer continues the history of its last item: image, counters, denoiser moments
    for (uint32_t i = 0; i < n; i++) {
        GpsRenderer *r = items[i].renderer;
        bool last = true;
        for (uint32_t k = i + 1; k < n && last; k++) last = items[k].renderer != r;
        if (!last) continue;
        const uint64_t X = (uint64_t)r->W * r->H;
        if (hipMemcpyAsync(r->accum, b->outs[i], X * sizeof(float4), hipMemcpyDeviceToDevice, b->stream) !=
                hipSuccess ||
            hipMemcpyAsync(r->stats, b->item_stats + (size_t)i * ST_WORDS, ST_WORDS * 8, hipMemcpyDeviceToDevice,
                           b->stream) != hipSuccess)
            rc = -1;
        if (r->features && (hipMemcpyAsync(r->mom, b->fbuf[i].mom, X * sizeof(float4), hipMemcpyDeviceToDevice,
                                           b->stream) != hipSuccess ||
                            hipMemcpyAsync(r->zmin, b->fbuf[i].zmin, X * sizeof(float), hipMemcpyDeviceToDevice,
                                           b->stream) != hipSuccess))
            rc = -1;
        r->accumulated = items[i].count;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) rc = -1;
    if (rc != 0) g_create_error = "fused batch copy-back failed";
    return rc;
}

int gps_batch_render(GpsBatch *b, const GpsBatchItem *items, uint32_t n, int use_graph) {
    b->gpu_ms = b->record_ms = 0.0;
    b->nodes = 0;
    if (n == 0) return 0;
    if (!items) {
        g_create_error = "null batch items";
        return -1;
    }
    if (hipSetDevice(b->device) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return -1;
    }
    // validate and settle every renderer before recording (allocation never happens while recording)
    for (uint32_t i = 0; i < n; i++) {
        GpsRenderer *r = items[i].renderer;
        if (!r || r->device != b->device) {
            g_create_error = "batch item without a renderer on the batch's device";
            return -1;
        }
        if (check_params(r, items[i].params) != 0) {
            g_create_error = r->err;
            return -1;
        }
        if (items[i].params->path == GPS_PATH_PARTICLE_SORTED) {
            g_create_error = "the sorted particle path cannot be recorded into a batch";
            return -1;
        }
        if (hipStreamSynchronize(r->stream) != hipSuccess) {
            g_create_error = "renderer stream sync failed";
            return -1;
        }
    }
    if (b->outs.size() < n) {
        b->outs.resize(n, nullptr);
        b->out_px.resize(n, 0);
    }
    if (b->item_stats_cap < n) {
        if (b->item_stats) (void)hipFree(b->item_stats);
        b->item_stats = nullptr;
        b->item_stats_cap = 0;
        if (hipMalloc(&b->item_stats, (size_t)n * ST_WORDS * sizeof(uint64_t)) != hipSuccess) {
            g_create_error = "out of device memory for batch counters";
            return -1;
        }
        b->item_stats_cap = n;
    }
    b->item_w.assign(n, 0);
    b->item_h.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        b->item_w[i] = items[i].renderer->W;
        b->item_h[i] = items[i].renderer->H;
    }
    for (uint32_t i = 0; i < n; i++) {
        GpsRenderer *r = items[i].renderer;
        uint64_t X = (uint64_t)r->W * r->H;
        if (b->out_px[i] < X) {
            if (b->outs[i]) (void)hipFree(b->outs[i]);
            b->outs[i] = nullptr;
            if (hipMalloc(&b->outs[i], X * sizeof(float4)) != hipSuccess) {
                g_create_error = "out of device memory for batch outputs";
                return -1;
            }
            b->out_px[i] = X;
        }
    }
    // every item on the particle path of a non-empty scene, without the debug hook: the fused kernels
    bool fusable = (use_graph & GPS_BATCH_PER_ITEM) == 0 && n <= (uint32_t)FUSED_MAX_ITEMS;
    for (uint32_t i = 0; i < n && fusable; i++) {
        const GpsRenderer *r = items[i].renderer;
        if (fused_mk(items[i].params) < 0 || r->n == 0 || (r->debug & GPS_DEBUG_KEEP_BUFFERS)) fusable = false;
    }
    if (fusable) {
        g_create_error.clear();
        return fused_batch_render(b, items, n, (use_graph & GPS_BATCH_GRAPH) != 0);
    }
    const bool graph_on = (use_graph & GPS_BATCH_GRAPH) != 0;
    auto t0 = std::chrono::steady_clock::now();
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (graph_on && hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        g_create_error = "hipStreamBeginCapture failed";
        return -1;
    }
    // lanes: one per renderer in order of first appearance, round robin
    std::vector<int> lane(n);
    std::vector<GpsRenderer *> seen;
    for (uint32_t i = 0; i < n; i++) {
        int g = -1;
        for (size_t k = 0; k < seen.size(); k++)
            if (seen[k] == items[i].renderer) g = (int)k;
        if (g < 0) {
            g = (int)seen.size();
            seen.push_back(items[i].renderer);
        }
        lane[i] = g % BATCH_LANES;
    }
    const int nl = seen.size() < (size_t)BATCH_LANES ? (int)seen.size() : BATCH_LANES;
    int rc = 0;
    if (hipEventRecord(b->fork, b->stream) != hipSuccess) rc = -1;
    for (int l = 0; l < nl && rc == 0; l++)
        if (hipStreamWaitEvent(b->lanes[l], b->fork, 0) != hipSuccess) rc = -1;
    for (uint32_t i = 0; i < n && rc == 0; i++) {
        const GpsBatchItem &it = items[i];
        GpsRenderer *r = it.renderer;
        hipStream_t st = b->lanes[lane[i]];
        uint64_t X = (uint64_t)r->W * r->H;
        launch_accum_clear(st, r->accum, X);
        launch_features_clear(st, r->mom, r->zmin, X);
        launch_stats_clear(st, r->stats, 0, ST_WORDS);
        r->accumulated = 0;
        rc = enqueue_ensembles(r, st, it.camera, it.params, it.first_seed, it.count);
        if (rc == 0) {
            launch_copy_f4(st, b->outs[i], r->accum, X);
            // the item's counters (ST_WORDS u64 = ST_WORDS / 2 float4)
            launch_copy_f4(st, reinterpret_cast<float4 *>(b->item_stats + (size_t)i * ST_WORDS),
                           reinterpret_cast<const float4 *>(r->stats), ST_WORDS / 2);
        } else {
            g_create_error = r->err;
        }
        r->enqueue_stream = r->stream;
    }
    // join every lane that was forked, also after a failure, so that a capture can end
    for (int l = 0; l < nl; l++)
        if (hipEventRecord(b->join[l], b->lanes[l]) != hipSuccess ||
            hipStreamWaitEvent(b->stream, b->join[l], 0) != hipSuccess)
            rc = -1;
    if (rc != 0 && g_create_error.empty()) g_create_error = "batch recording failed";
    b->nodes = 0;
    if (graph_on) {
        if (hipStreamEndCapture(b->stream, &graph) != hipSuccess) rc = -1;
        if (rc == 0 && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = -1;
        size_t nn = 0;
        if (rc == 0 && hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess) b->nodes = nn;
        if (rc != 0 && g_create_error.empty()) g_create_error = "graph capture failed";
    }
    auto t1 = std::chrono::steady_clock::now();
    b->record_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == 0 && graph_on) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, b->stream);
        if (hipGraphLaunch(exec, b->stream) != hipSuccess) rc = -1;
        (void)hipEventRecord(e1, b->stream);
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) rc = -1;
    if (e0) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        b->gpu_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        b->gpu_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

int gps_batch_read(GpsBatch *b, uint32_t item, float *out) {
    // the last render's item size: an output buffer keeps the largest size it has held
    if (item >= b->item_w.size() || item >= b->outs.size() || !b->outs[item]) return -1;
    const uint64_t X = (uint64_t)b->item_w[item] * b->item_h[item];
    if (hipMemcpy(out, b->outs[item], X * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

int gps_batch_composite(GpsBatch *b, uint32_t item, const GpsRenderParams *p, uint32_t out_w, uint32_t out_h,
                        uint8_t *rgba8) {
    if (item >= b->item_w.size() || !b->outs[item] || out_w == 0 || out_h == 0) return -1;
    if (hipSetDevice(b->device) != hipSuccess) return -1;
    const uint64_t px = (uint64_t)out_w * out_h;
    if (px > b->composite_cap) {
        if (b->composite_buf) (void)hipFree(b->composite_buf);
        b->composite_buf = nullptr;
        b->composite_cap = 0;
        if (hipMalloc(&b->composite_buf, px * 4) != hipSuccess) return -1;
        b->composite_cap = px;
    }
    launch_composite(b->stream, b->outs[item], b->item_w[item], b->item_h[item], p->exposure,
                     p->srgb_output ? 1u : 0u, out_w, out_h, b->composite_buf);
    if (hipMemcpyAsync(rgba8, b->composite_buf, px * 4, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return -1;
    return 0;
}

int gps_batch_read_stats(GpsBatch *b, uint32_t item, GpsStats *out) {
    if (item >= b->item_stats_cap || item >= b->outs.size() || !b->outs[item]) return -1;
    uint64_t w[ST_WORDS];
    if (hipMemcpy(w, b->item_stats + (size_t)item * ST_WORDS, sizeof w, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    stats_from_words(w, out);
    return 0;
}

void gps_batch_timing(const GpsBatch *b, double *gpu_ms, double *record_ms, uint64_t *graph_nodes) {
    if (gpu_ms) *gpu_ms = b->gpu_ms;
    if (record_ms) *record_ms = b->record_ms;
    if (graph_nodes) *graph_nodes = b->nodes;
}

// -------------------------------------------------------------- tuning
void gps_tuning_get(GpsTuning *out) {
    out->particle_waves_per_cu = g_tuning.particle_waves_per_cu;
    out->particles_per_thread = g_tuning.particles_per_thread;
    out->sort_items_per_thread = g_tuning.sort_items_per_thread;
    out->batch_ensembles = g_tuning.batch_ensembles;
    out->depth_window_mib = g_tuning.depth_window_mib;
    out->count_only = g_tuning.count_only;
    out->particle_block = g_tuning.particle_block;
    out->sort_digit_bits = g_tuning.sort_digit_bits;
}

int gps_tuning_set(const GpsTuning *t) {
    uint32_t ppt = t->particles_per_thread;
    if (!(ppt == 1 || ppt == 2 || ppt == 4 || ppt == 8)) return -1;
    const uint32_t si = t->sort_items_per_thread;
    if (!(si == 8 || si == 16 || si == 32 || si == 64)) return -1;
    if (t->particle_waves_per_cu < 1 || t->particle_waves_per_cu > 1024) return -1;
    if (t->batch_ensembles > MAX_BATCH) return -1;
    if (t->depth_window_mib < 1 || t->depth_window_mib > (1u << 20)) return -1;
    if (t->count_only > 1) return -1;
    const uint32_t pb = t->particle_block;
    if (!(pb == 128 || pb == 256 || pb == 512)) return -1;
    if (t->sort_digit_bits < 4 || t->sort_digit_bits > 8) return -1;
    g_tuning.particle_waves_per_cu = t->particle_waves_per_cu;
    g_tuning.particles_per_thread = ppt;
    g_tuning.sort_items_per_thread = t->sort_items_per_thread;
    g_tuning.batch_ensembles = t->batch_ensembles;
    g_tuning.depth_window_mib = t->depth_window_mib;
    g_tuning.count_only = t->count_only;
    g_tuning.particle_block = pb;
    g_tuning.sort_digit_bits = t->sort_digit_bits;
    return 0;
}

// Times each candidate tuning on `trial_ensembles` ensembles of this renderer's scene and view (seeds from
// first_seed) and keeps the fastest; the accumulation buffer, the counters and `submitted` are restored.
int gps_renderer_autotune(GpsRenderer *r, const GpsCamera *cam, const GpsRenderParams *p, uint32_t first_seed,
                          uint32_t trial_ensembles, GpsTuning *best_out, double *best_ms_out) {
    if (check_params(r, p) != 0) return -1;
    if (trial_ensembles == 0) trial_ensembles = 1;
    const uint64_t X = (uint64_t)r->W * r->H;
    float4 *acc_save = nullptr, *mom_save = nullptr;
    uint64_t *st_save = nullptr;
    float *z_save = nullptr;
    HIP_TRY(r, hipMalloc(&acc_save, X * sizeof(float4)));
    if (hipMalloc(&st_save, ST_WORDS * 8) != hipSuccess ||
        (r->features && (hipMalloc(&mom_save, X * sizeof(float4)) != hipSuccess ||
                         hipMalloc(&z_save, X * sizeof(float)) != hipSuccess))) {
        dfree(acc_save);
        dfree(st_save);
        dfree(mom_save);
        dfree(z_save);
        return fail(r, "hipMalloc failed");
    }
    HIP_TRY(r, hipMemcpyAsync(acc_save, r->accum, X * sizeof(float4), hipMemcpyDeviceToDevice, r->stream));
    HIP_TRY(r, hipMemcpyAsync(st_save, r->stats, ST_WORDS * 8, hipMemcpyDeviceToDevice, r->stream));
    if (r->features) {
        HIP_TRY(r, hipMemcpyAsync(mom_save, r->mom, X * sizeof(float4), hipMemcpyDeviceToDevice, r->stream));
        HIP_TRY(r, hipMemcpyAsync(z_save, r->zmin, X * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
    }
    const uint32_t submitted = r->accumulated;
    const GpsTuningDev start = g_tuning;
    GpsTuningDev best = start;
    double best_ms = 1e30;
    const uint32_t waves[] = {16u, 32u, 64u, 128u, 256u};
    const uint32_t ppts[] = {1u, 2u, 4u, 8u};
    const uint32_t items[] = {8u, 16u, 32u, 64u};
    const uint32_t blocks[] = {128u, 256u, 512u};
    const bool sorted_path = p->path == GPS_PATH_PARTICLE_SORTED;
    hipEvent_t e0, e1;
    HIP_TRY(r, hipEventCreate(&e0));
    HIP_TRY(r, hipEventCreate(&e1));
    int rc = 0;
    // one untimed call first so that lazily sized buffers exist before timing
    rc = gps_renderer_render_ensembles(r, cam, p, first_seed, 1);
    for (uint32_t bl : blocks)
        for (uint32_t wv : waves)
            for (uint32_t ppt : ppts)
                for (uint32_t it : items) {
                    if (rc != 0) break;
                    if (!sorted_path && it != start.sort_items_per_thread) continue;
                    // the same threads per CU: waves x 256 threads of grid, whatever the workgroup size
                    g_tuning.particle_block = bl;
                    g_tuning.particle_waves_per_cu = wv * 256u / bl;
                    g_tuning.particles_per_thread = ppt;
                    g_tuning.sort_items_per_thread = it;
                    (void)hipEventRecord(e0, r->stream);
                    rc = gps_renderer_render_ensembles(r, cam, p, first_seed, trial_ensembles);
                    (void)hipEventRecord(e1, r->stream);
                    (void)hipEventSynchronize(e1);
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (rc == 0 && ms < best_ms) {
                        best_ms = ms;
                        best = g_tuning;
                    }
                }
    double best_ms_per = best_ms / trial_ensembles;
    // sorted path: the radix digit width of the 64-bit sort (narrower digits: cheaper passes, more of them)
    if (rc == 0 && sorted_path) {
        for (uint32_t bits = 4; bits <= 8 && rc == 0; bits++) {
            if (bits == best.sort_digit_bits) continue;
            g_tuning = best;
            g_tuning.sort_digit_bits = bits;
            (void)hipEventRecord(e0, r->stream);
            rc = gps_renderer_render_ensembles(r, cam, p, first_seed, trial_ensembles);
            (void)hipEventRecord(e1, r->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double per = (double)ms / trial_ensembles;
            if (rc == 0 && per < best_ms_per) {
                best_ms_per = per;
                best = g_tuning;
            }
        }
    }
    // second stage: ensembles per particle launch, through the depth window
    if (rc == 0 && !sorted_path && start.batch_ensembles == 0) {
        const uint32_t windows[] = {8u, 16u, 32u, 64u, 128u, 256u, 512u, 4096u};
        uint32_t last_b = 0;
        for (uint32_t wmib : windows) {
            if (rc != 0) break;
            g_tuning = best;
            g_tuning.depth_window_mib = wmib;
            const uint32_t b = ensembles_per_launch(r);
            if (b == last_b) continue;  // same launches as the previous window
            last_b = b;
            const uint32_t trial = trial_ensembles > 2 * b ? trial_ensembles : 2 * b;
            (void)hipEventRecord(e0, r->stream);
            rc = gps_renderer_render_ensembles(r, cam, p, first_seed, trial);
            (void)hipEventRecord(e1, r->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double per = (double)ms / trial;
            if (rc == 0 && per < best_ms_per) {
                best_ms_per = per;
                best = g_tuning;
            }
        }
    }
    g_tuning = rc == 0 ? best : start;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipMemcpyAsync(r->accum, acc_save, X * sizeof(float4), hipMemcpyDeviceToDevice, r->stream);
    (void)hipMemcpyAsync(r->stats, st_save, ST_WORDS * 8, hipMemcpyDeviceToDevice, r->stream);
    if (r->features) {
        (void)hipMemcpyAsync(r->mom, mom_save, X * sizeof(float4), hipMemcpyDeviceToDevice, r->stream);
        (void)hipMemcpyAsync(r->zmin, z_save, X * sizeof(float), hipMemcpyDeviceToDevice, r->stream);
    }
    (void)hipStreamSynchronize(r->stream);
    dfree(acc_save);
    dfree(st_save);
    dfree(mom_save);
    dfree(z_save);
    r->accumulated = submitted;
    if (rc != 0) return -1;
    if (best_out) gps_tuning_get(best_out);
    if (best_ms_out) *best_ms_out = best_ms_per;
    return 0;
}

// -------------------------------------------------------------- phase-0 unit entry points
int gps_phase0_atomic(int device, uint32_t seed, uint32_t pixels, uint32_t samples, uint32_t *depth, uint32_t *color) {
    if (pixels == 0) return -1;
    if (hipSetDevice(device) != hipSuccess) return -1;
    uint64_t *buf = nullptr;
    if (hipMalloc(&buf, (uint64_t)pixels * 8) != hipSuccess) return -1;
    launch_fill_u64(nullptr, buf, pixels, EMPTY64);
    launch_phase0(nullptr, seed, pixels, samples, buf);
    std::vector<uint64_t> h(pixels);
    int rc = hipMemcpy(h.data(), buf, (uint64_t)pixels * 8, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
    (void)hipFree(buf);
    for (uint32_t i = 0; i < pixels && rc == 0; i++) {
        depth[i] = (uint32_t)(h[i] >> 32);
        color[i] = (uint32_t)h[i];
    }
    return rc;
}

int gps_scan_exclusive_u32(int device, const uint32_t *in, uint32_t *out, uint64_t n) {
    if (n == 0) return 0;
    if (hipSetDevice(device) != hipSuccess) return -1;
    uint32_t tiles = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    uint32_t *din = nullptr, *dout = nullptr, *ticket = nullptr;
    uint64_t *status = nullptr;
    int rc = 0;
    if (hipMalloc(&din, n * 4) != hipSuccess || hipMalloc(&dout, n * 4) != hipSuccess ||
        hipMalloc(&ticket, 4) != hipSuccess || hipMalloc(&status, (uint64_t)tiles * 8) != hipSuccess)
        rc = -1;
    if (rc == 0 && (hipMemcpy(din, in, n * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemset(ticket, 0, 4) != hipSuccess || hipMemset(status, 0, (uint64_t)tiles * 8) != hipSuccess))
        rc = -1;
    if (rc == 0) {
        launch_scan_u32(nullptr, din, dout, n, ticket, status, tiles, 1u);
        if (hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    }
    dfree(din);
    dfree(dout);
    dfree(ticket);
    dfree(status);
    return rc;
}

int gps_math_eval(int device, int fn, const float *x, const float *y, float *out, uint64_t n) {
    if (n == 0) return 0;
    if (hipSetDevice(device) != hipSuccess) return -1;
    float *dx = nullptr, *dy = nullptr, *dout = nullptr;
    int rc = 0;
    if (hipMalloc(&dx, n * 4) != hipSuccess || hipMalloc(&dy, n * 4) != hipSuccess ||
        hipMalloc(&dout, n * 4) != hipSuccess)
        rc = -1;
    if (rc == 0 && hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    if (rc == 0 && y && hipMemcpy(dy, y, n * 4, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    if (rc == 0) {
        launch_math(nullptr, fn, dx, dy, dout, n);
        if (hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    }
    dfree(dx);
    dfree(dy);
    dfree(dout);
    return rc;
}

}  // extern "C"
