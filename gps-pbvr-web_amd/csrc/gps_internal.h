// Internal definitions shared by the kernel and host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Per-view constants of one render call, passed to every kernel by value.
// Field meaning follows the reference's FrameUniform (crates/gps-render/src/gps.rs:63-76,
// packed by upload_frame, gps.rs:812-856); the ensemble seed is not part of it.
struct GpsFrame {
    float rot[3][4];        // rows of view_rot
    float pos[4];           // camera position, w = low-pass variance (px^2)
    float intr[4];          // fx, fy, cx, cy
    uint32_t W, H, spp_side, N;
    uint32_t cap, sh_stride, exact_radius, sh_degree;
    uint32_t method, calib, flags, pad0;
    float density, near_z, base_k, ref_len;
    float object_depth, jitter, pad2, pad3;   // jitter = sqrt(max(pos.w, 0)), the same for every particle
    float bg[4];
    // (float)W, (float)H, (float)spp_side, converted on the host (pack_frame): the particle kernels' emitters take
    // them as uniform operands, where a conversion in the kernel held registers that the seven-wave occupancy
    // target then spilled inside the particle loop
    float wf, hf, sidef, pad4;
};

enum : uint32_t {
    M_GPS = 0, M_PROPORTIONAL = 1, M_EXTINCTION = 2, M_VIEW_CONDITIONED = 3, M_OCCUPANCY = 4,
    FLAG_RADIAL = 1, FLAG_CENTRE = 2, FLAG_JITTER = 4, FLAG_CULL = 8,
    FLAG_FAST = 16,   // PBVR: count the particles of off-screen splats on the count-only path
};

// per-splat count source written by the view kernel (kind in the low bits of .y)
enum : uint32_t { KIND_NONE = 0, KIND_POISSON = 1, KIND_FIXED = 2, KIND_UNDERCOVERED = 4 };

// device statistics slots (u64 each)
enum : int {
    ST_POINTS = 0, ST_TRUNC = 1, ST_DROPPED = 2, ST_VISIBLE = 3, ST_ORPHAN = 4, ST_CAND = 5, ST_UNDER = 6,
    ST_SKIPPED = 7, ST_ENSEMBLES = 8, ST_PROBE = 9,
    // high-water marks of the particle kernel's count-only LDS buffers since the last stats reset: a wave's
    // emit_pbvr queue and a wave's count_only work buffer (capacity PARTICLE_WAVE_CAP each)
    ST_QPEAK = 10, ST_WPEAK = 11,
    // particles of the particle kernel that reached a subpixel (the packed minimum or a sorted key on screen)
    ST_HITS = 12,
    ST_WORDS = 16,
};

constexpr int VIEW_VEC4 = 6;          // float4 per splat in the view cache
constexpr uint32_t PARTICLE_WAVE_CAP = 128;   // count-only work records / emit_pbvr entries per wave (LDS)
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int MAX_BATCH = 64;         // ensembles per batch (one depth|id buffer each)
constexpr uint64_t EMPTY64 = ~0ull;

// process-wide launch tuning (include/gps_hip.h GpsTuning); results never depend on it
struct GpsTuningDev {
    uint32_t particle_waves_per_cu;   // particle grid = CUs * this workgroups (chunks are grid-strided)
    uint32_t particles_per_thread;    // 1, 2, 4 or 8
    uint32_t sort_items_per_thread;   // 8 or 16
    uint32_t batch_ensembles;         // ensembles per particle launch, 0 = from depth_window_mib
    uint32_t depth_window_mib;        // depth|id bytes one launch may span when batch_ensembles is 0
    uint32_t count_only;              // 1: count-only path for the particles of off-screen PBVR splats
    uint32_t particle_block;          // particle kernel workgroup size: 128, 256 or 512 threads
    uint32_t sort_digit_bits;         // radix digit width of the 64-bit sort, 4..8
};
GpsTuningDev gps_tuning_current();

struct BatchState {
    uint32_t *tickets;      // [MAX_BATCH] scan tile tickets (reset by the last tile)
    uint64_t *status;       // [MAX_BATCH * tiles] decoupled look-back words
    uint32_t *totals;       // [MAX_BATCH] particle total of each ensemble (0 if skipped)
    uint32_t *skip;         // [MAX_BATCH] 1 if the ensemble overflowed u32
};

// One item of a fused batch (gps_batch_render): a view of a renderer's scene, rendered with buffers the batch owns
// for it.  Every launch of a fused batch covers all items; a round renders up to `round_ens` ensembles of each item.
struct FusedItem {
    GpsFrame f;
    const float4 *splats;
    const float *sh;
    const float *lut;
    float4 *view;           // [N] x 6 float4
    uint2 *cinfo;           // [N]
    float4 *colour;         // [N]
    uint32_t *perm;         // [N + 3] count-only splat order, or null
    uint32_t *offsets;      // [round_ens][N]
    uint64_t *depth;        // [round_ens][sub_count], all ones between rounds
    BatchState bs;          // [MAX_BATCH] tickets, totals, skip; [round_ens][tiles] status
    float4 *accum;          // [X] the item's image
    float4 *mom;            // [X] denoiser moments, or null
    float *zmin;            // [X]
    uint64_t *stats;        // [ST_WORDS]
    uint64_t sub_count;
    uint32_t tiles;         // scan tiles of N
    uint32_t first_seed;
    uint32_t mk;            // particle emitter: 0 GPS, 1 PBVR, 2 occupancy
};
constexpr uint32_t FUSED_NONE = 0xFFFFFFFFu;
constexpr int FUSED_MAX_ITEMS = 1024;

// kernel launchers (gps_kernels.hip)
// perm (optional, [N + 3]): the view kernel writes the splat order with the count-only splats first into
// perm[0, N) and counts them in the 8-byte word at perm + perm_counter(N) (zeroed before); the count scan and
// the packed-minimum particle kernel then work in that order (results do not depend on the order)
__host__ __device__ inline uint32_t perm_counter(uint32_t n) { return (n + 1u) & ~1u; }
void launch_view(hipStream_t s, const GpsFrame &f, const float4 *splats, const float *sh, float4 *view,
                 uint2 *cinfo, float4 *colour, uint32_t *perm = nullptr);
void launch_count_scan(hipStream_t s, const GpsFrame &f, const uint2 *cinfo, uint32_t first_seed, uint32_t nens,
                       uint32_t *offsets, const BatchState &bs, uint32_t tiles, uint32_t epoch, uint64_t *stats,
                       const uint32_t *perm = nullptr);
void launch_particles(hipStream_t s, const GpsFrame &f, const float4 *view, const uint32_t *offsets,
                      const BatchState &bs, uint32_t nens, uint32_t first_seed, uint64_t *depth, uint64_t sub_count,
                      const float *lut, uint64_t *stats, int grid, uint64_t *pkeys = nullptr,
                      uint32_t *pidx = nullptr, int ppt = 4, const uint32_t *perm = nullptr, int block = 256);
void launch_resolve(hipStream_t s, const GpsFrame &f, uint64_t *depth, uint64_t sub_count, const uint32_t *skip,
                    uint32_t nens, const float4 *colour, float4 *accum, int clear, uint64_t *stats,
                    float4 *mom = nullptr, float *zmin = nullptr);
void launch_features_clear(hipStream_t s, float4 *mom, float *zmin, uint64_t n);
void launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t n, uint64_t v);
void launch_splat_pass(hipStream_t s, const GpsFrame &f, const float4 *view, const uint2 *cinfo, uint32_t seed,
                       uint64_t *depth, const float *lut, uint64_t *stats);
void launch_composite(hipStream_t s, const float4 *accum, uint32_t W, uint32_t H, float exposure, uint32_t gamma_on,
                      uint32_t ow, uint32_t oh, uint32_t *out);
void launch_zero_u32(hipStream_t s, uint32_t *p, uint64_t n);
void launch_stats_add(hipStream_t s, uint64_t *stats, int slot, uint64_t v);
void launch_stats_clear(hipStream_t s, uint64_t *stats, int first, int count);
void launch_accum_clear(hipStream_t s, float4 *accum, uint64_t n);
void launch_copy_f4(hipStream_t s, float4 *dst, const float4 *src, uint64_t n);
void launch_phase0(hipStream_t s, uint32_t seed, uint32_t pixels, uint32_t samples, uint64_t *buf);
void launch_scan_u32(hipStream_t s, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *ticket, uint64_t *status,
                     uint32_t tiles, uint32_t epoch);
void launch_math(hipStream_t s, int fn, const float *x, const float *y, float *out, uint64_t n);
// fused batch launches (gps_kernels.hip); tables are device arrays built by the host at record time
void launch_fused_clear(hipStream_t s, const FusedItem *items, uint32_t nitems, uint32_t max_x);
void launch_fused_view(hipStream_t s, const FusedItem *items, const uint32_t *blk_item, const uint32_t *item_blk0,
                       uint32_t blocks);
void launch_fused_count_scan(hipStream_t s, const FusedItem *items, const uint32_t *blk_pair, uint32_t blocks,
                             uint32_t ens0, uint32_t epoch);
void launch_fused_assign(hipStream_t s, const FusedItem *items, uint32_t nitems, const uint32_t *nens, uint32_t mk,
                         uint32_t chunk, uint32_t grid, uint32_t *blk_item, uint32_t *blk_rank, uint32_t *item_nblk);
void launch_fused_particles(hipStream_t s, const FusedItem *items, uint32_t mk, const uint32_t *nens, uint32_t ens0,
                            const uint32_t *blk_item, const uint32_t *blk_rank, const uint32_t *item_nblk, int grid,
                            int ppt, int block);
void launch_fused_resolve(hipStream_t s, const FusedItem *items, const uint32_t *nens, const uint32_t *blk_item,
                          const uint32_t *item_blk0, uint32_t blocks);

// device buffers of a renderer for the denoiser: accumulation, moments and nearest depth (null when the
// renderer's features are off)
struct GpsRenderer;
void renderer_device_buffers(GpsRenderer *r, const float4 **accum, const float4 **mom, const float **zmin);
