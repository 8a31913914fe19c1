// Depth-sorted alpha path (crates/gps-render/src/sorted.rs; shaders/gps_splat.wgsl:736-786,
// gps_sort.wgsl, gps_sorted_draw.wgsl).
//
// CDNA has no rasterizer, so the reference's instanced-quad draw with hardware SrcAlpha /
// OneMinusSrcAlpha blending becomes compute:
//   k_sp_project   per splat: cull, GPS demand probe, quad axes, SH colour, 24-bit far->near key
//   radix sort     stable LSD, digits of equal width of at most 8 bits: per-tile digit histograms
//                  (digit-major), a decoupled look-back scan over them, then a stable scatter ranked with
//                  wave64 ballots
//   binning        each visible splat (in sorted order) emits one (tile, splat) pair per 16x16 tile its
//                  quad touches, balanced by pairs; a stable sort by tile keeps the back-to-front order
//                  inside each tile; the tiles are then ordered longest list first
//   k_sp_blend     one workgroup per 8x8 quadrant of a tile: the splats that may reach the quadrant (binning
//                  box, ellipse bound) are kept, four waves evaluate their alphas, one runs the back-to-front
//                  blend at pixel centres in f32, written to the accumulation buffer with count 1 (sp_resolve)
#include <string>
#include <vector>

#include "gps_devmath.h"
#include "gps_internal.h"

namespace {

using gpsm::clamp;
using gpsm::f2u;
using gpsm::fmax;
using gpsm::fmin;

constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int BIN = 16;                      // screen tile side in pixels
#ifndef GPS_SP_SCAN_ITEMS
#define GPS_SP_SCAN_ITEMS 8
#endif
constexpr int SP_SCAN_ITEMS = GPS_SP_SCAN_ITEMS;   // items per thread of k_rs_scan (256 threads per tile)
constexpr uint32_t SP_SCAN_TILE = 256u * SP_SCAN_ITEMS;
constexpr uint32_t SP_KEY_INVALID = 0xFFFFFFu;
constexpr float SP_MIN_ALPHA = 0.00392156862745098f;
constexpr float PI_F = 3.14159265358979f;
constexpr float DILOG_ONE = 1.64493406684823f;

// ------------------------------------------------------------------ arithmetic shared with gps_kernels.hip
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

GPS_DEV float dot3(const float *r, float x, float y, float z) { return r[0] * x + r[1] * y + r[2] * z; }

struct Proj {
    bool ok;
    float mx, my, a, b, c, depth;
};

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

// gps_splat.wgsl:210-251
GPS_DEV float4 splat_colour(const GpsFrame &f, uint32_t id, float4 g0, float4 g3, const float *__restrict__ sh) {
    float res[3] = {g3.x, g3.y, g3.z};
    uint32_t deg = f.sh_degree;
    if (deg == 0u) return make_float4(fmax(res[0], 0.0f), fmax(res[1], 0.0f), fmax(res[2], 0.0f), 0.0f);
    float px = g0.x - f.pos[0], py = g0.y - f.pos[1], pz = g0.z - f.pos[2];
    float len = gpsm::sqrt(px * px + py * py + pz * pz);
    float x = px / len, y = py / len, z = pz / len;
    uint32_t r = f.sh_stride;
    uint32_t o = id * 3u * r;
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

// ------------------------------------------------------------------ k_sp_project (gps_splat.wgsl:736-777)
__global__ __launch_bounds__(256) void k_sp_project(GpsFrame f, const float4 *__restrict__ splats,
                                                    const float *__restrict__ sh, float4 *__restrict__ sp,
                                                    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                    uint4 *__restrict__ part, int probe_only) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t demand = 0, vis = 0;
    if (i < f.N) {
        if (!probe_only) {
            keys[i] = SP_KEY_INVALID;
            vals[i] = i;
        }
        float4 g0 = splats[4u * i], g1 = splats[4u * i + 1u], g2 = splats[4u * i + 2u];
        Proj pr = project(f, g0, g1, g2);
        float o = clamp(g0.w, 0.0f, 1.0f);
        float ex = 5.9f * gpsm::sqrt(pr.a), ey = 5.9f * gpsm::sqrt(pr.c);
        bool off = pr.mx + ex < 0.0f || pr.mx - ex >= (float)f.W || pr.my + ey < 0.0f || pr.my - ey >= (float)f.H;
        if (pr.ok && o > 0.0f && !off) {
            uint32_t spp = f.spp_side * f.spp_side;
            float det0 = pr.a * pr.c - pr.b * pr.b;
            float en = f.density * (float)spp * 2.0f * PI_F * gpsm::sqrt(det0) * dilog(o);
            if (en > 0.0f) demand = f2u(fmin(en, 4.0e9f));
            if (!probe_only && !(o < SP_MIN_ALPHA)) {
                float k = gpsm::sqrt(2.0f * gpsm::log(o / SP_MIN_ALPHA));
                float kx = k * gpsm::sqrt(pr.a), ky = k * gpsm::sqrt(pr.c);
                if (!(pr.mx + kx < 0.0f || pr.mx - kx >= (float)f.W || pr.my + ky < 0.0f ||
                      pr.my - ky >= (float)f.H)) {
                    float tr2 = 0.5f * (pr.a + pr.c);
                    float disc = gpsm::sqrt(fmax(tr2 * tr2 - (pr.a * pr.c - pr.b * pr.b), 0.0f));
                    float l1 = tr2 + disc, l2 = fmax(tr2 - disc, 1e-12f);
                    float v1x = 1.0f, v1y = 0.0f;
                    float absb = pr.b < 0.0f ? -pr.b : pr.b;
                    if (absb > 1e-12f) {
                        float ax = l1 - pr.c, ay = pr.b;
                        float len = gpsm::sqrt(ax * ax + ay * ay);
                        v1x = ax / len;
                        v1y = ay / len;
                    } else if (pr.a < pr.c) {
                        v1x = 0.0f;
                        v1y = 1.0f;
                    }
                    float v2x = -v1y, v2y = v1x;
                    float4 col = splat_colour(f, i, g0, splats[4u * i + 3u], sh);
                    float s1 = gpsm::sqrt(l1), s2 = gpsm::sqrt(l2);
                    sp[3u * i] = make_float4(pr.mx, pr.my, o, k);
                    sp[3u * i + 1u] = make_float4(v1x * s1, v1y * s1, v2x * s2, v2y * s2);
                    sp[3u * i + 2u] = make_float4(col.x, col.y, col.z, 0.0f);
                    keys[i] = 0xFFFFFEu - (gpsm::bits(pr.depth) >> 8u);
                    vis = 1;
                }
            }
        }
    }
    // the block's demand and visible count as one partial (k_sp_totals adds them up): an atomic per wave on the
    // same two words would serialize thousands of waves in one L2 channel
    __shared__ uint64_t s_dem[4];
    __shared__ uint32_t s_vis[4];
    uint64_t d64 = demand;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        d64 += __shfl_xor(d64, off, 64);
        vis += __shfl_xor(vis, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_dem[threadIdx.x >> 6] = d64;
        s_vis[threadIdx.x >> 6] = vis;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t d = s_dem[0] + s_dem[1] + s_dem[2] + s_dem[3];
        part[blockIdx.x] = make_uint4((uint32_t)d, (uint32_t)(d >> 32), s_vis[0] + s_vis[1] + s_vis[2] + s_vis[3], 0u);
    }
}

// the frame's probe demand (stats[ST_PROBE]) and visible splat count from k_sp_project's per-block partials
constexpr int TOTALS_THREADS = 1024;
__global__ __launch_bounds__(TOTALS_THREADS) void k_sp_totals(const uint4 *__restrict__ part, uint32_t nparts,
                                                              uint64_t *__restrict__ stats,
                                                              uint32_t *__restrict__ visible, int probe_only) {
    __shared__ uint64_t s_dem[TOTALS_THREADS / 64];
    __shared__ uint32_t s_vis[TOTALS_THREADS / 64];
    uint64_t d = 0;
    uint32_t v = 0;
    for (uint32_t b = threadIdx.x; b < nparts; b += TOTALS_THREADS) {
        const uint4 p = part[b];
        d += ((uint64_t)p.y << 32) | p.x;
        v += p.z;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        d += __shfl_xor(d, off, 64);
        v += __shfl_xor(v, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_dem[threadIdx.x >> 6] = d;
        s_vis[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d = 0;
        v = 0;
        for (int w = 0; w < TOTALS_THREADS / 64; w++) {
            d += s_dem[w];
            v += s_vis[w];
        }
        stats[ST_PROBE] = d;
        *visible = probe_only ? 0u : v;
    }
}

// ------------------------------------------------------------------ radix sort
// RB-bit digits (4 to 8): a pass sorts by (key >> shift) & (2^RB - 1)
template <int RB>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint32_t *__restrict__ keys, uint32_t n, uint32_t shift,
                                                      uint32_t ntiles, uint32_t *__restrict__ hist) {
    constexpr uint32_t NB = 1u << RB;
    static_assert(RB >= 4 && RB <= 8 && (int)NB <= RS_BLOCK, "digit width");
    __shared__ uint32_t h[NB];
    if (threadIdx.x < NB) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        uint32_t i = tile * RS_TILE + j * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (NB - 1u)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NB) hist[threadIdx.x * ntiles + tile] = h[threadIdx.x];
}

// Stable scatter.  Wave w owns items [tile*TILE + w*64*ITEMS, +64*ITEMS) in (j, lane) order, so ranks
// follow the input order: earlier waves, then earlier j, then lower lanes.
template <int RB>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint32_t *__restrict__ kin,
                                                         const uint32_t *__restrict__ vin,
                                                         uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                         uint32_t n, uint32_t shift, uint32_t ntiles,
                                                         const uint32_t *__restrict__ hist) {
    constexpr int NW = RS_BLOCK / 64;
    constexpr uint32_t NB = 1u << RB;   // digits; NB itself marks an item beyond n
    __shared__ uint32_t wh[NW][NB + 1];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t d = lane; d < NB + 1u; d += 64) wh[w][d] = 0;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t key[RS_ITEMS], val[RS_ITEMS], dig[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        uint32_t i = tile * RS_TILE + w * (64 * RS_ITEMS) + j * 64 + lane;
        bool valid = i < n;
        key[j] = valid ? kin[i] : 0u;
        val[j] = valid && vin ? vin[i] : 0u;
        uint32_t d = valid ? ((key[j] >> shift) & (NB - 1u)) : NB;
        dig[j] = d;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b <= RB; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t before = wh[w][d];
        __builtin_amdgcn_wave_barrier();
        rank[j] = before + (uint32_t)__popcll(peers & lt);
        if ((peers & lt) == 0ull) wh[w][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    if (threadIdx.x < NB) {
        uint32_t run = 0;
        for (int ww = 0; ww < NW; ww++) {
            uint32_t c = wh[ww][threadIdx.x];
            wh[ww][threadIdx.x] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        uint32_t d = dig[j];
        if (d < NB) {
            uint32_t pos = hist[d * ntiles + tile] + wh[w][d] + rank[j];
            kout[pos] = key[j];
            if (vout) vout[pos] = val[j];
        }
    }
}

// exclusive scan of the digit-major histogram: same look-back scheme as k_count_scan
GPS_DEV uint64_t sat_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    const uint64_t m = (1ull << 36) - 1ull;
    return s > m ? m : s;
}

// In-place exclusive scan of data[0..n) (decoupled look-back over 2048-item tiles, sums saturating at 2^36 - 1);
// with total set, the inclusive sum of all n items is written there too.
__global__ __launch_bounds__(256) void k_rs_scan(uint32_t *__restrict__ data, uint32_t n, uint32_t *ticket,
                                                 uint64_t *status, uint32_t tiles, uint32_t epoch,
                                                 uint64_t *__restrict__ total) {
    constexpr int ITEMS = SP_SCAN_ITEMS, TILE = 256 * ITEMS;
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_wave[4];
    __shared__ uint64_t s_prefix;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * TILE + threadIdx.x * ITEMS;
    uint32_t v[ITEMS];
    uint64_t tsum = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        v[j] = base + j < n ? data[base + j] : 0u;
        tsum += v[j];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint64_t wave_base = 0, agg = 0;
    for (int ww = 0; ww < 4; ww++) {
        if (ww < wave) wave_base += s_wave[ww];
        agg += s_wave[ww];
    }
    if (threadIdx.x == 0) {
        const uint64_t m = (1ull << 36) - 1ull;
        const uint64_t tag = (uint64_t)(epoch & ((1u << 26) - 1u)) << 38;
        uint64_t a = agg > m ? m : agg, excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&status[0], tag | (2ull << 36) | a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&status[tile], tag | (1ull << 36) | a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int32_t t = (int32_t)tile - 1;
            for (;;) {
                uint64_t wv;
                for (;;) {
                    wv = __hip_atomic_load(&status[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((uint32_t)(wv >> 38) == (epoch & ((1u << 26) - 1u)) && ((wv >> 36) & 3ull) != 0ull) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                excl = sat_add(excl, wv & m);
                if (((wv >> 36) & 3ull) == 2ull) break;
                t--;
            }
            __hip_atomic_store(&status[tile], tag | (2ull << 36) | sat_add(excl, a), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = excl;
    }
    __syncthreads();
    uint64_t run = s_prefix + wave_base + inc - tsum;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        if (base + j < n) data[base + j] = (uint32_t)run;
        run += v[j];
    }
    if (total && base < n && n - base <= (uint32_t)ITEMS) *total = run;   // the thread holding item n - 1
    if (tile == tiles - 1u && threadIdx.x == 0) *ticket = 0u;
}

// ------------------------------------------------------------------ binning
GPS_DEV void quad_tiles(const float4 *sp, uint32_t id, uint32_t W, uint32_t H, uint32_t &tx0, uint32_t &ty0,
                        uint32_t &tx1, uint32_t &ty1) {
    float4 a = sp[3u * id], b = sp[3u * id + 1u];
    float ex = a.w * (fabsf(b.x) + fabsf(b.z)) + 1.0f, ey = a.w * (fabsf(b.y) + fabsf(b.w)) + 1.0f;
    float x0 = floorf(a.x - ex), x1 = ceilf(a.x + ex), y0 = floorf(a.y - ey), y1 = ceilf(a.y + ey);
    x0 = clamp(x0, 0.0f, (float)(W - 1));
    x1 = clamp(x1, 0.0f, (float)(W - 1));
    y0 = clamp(y0, 0.0f, (float)(H - 1));
    y1 = clamp(y1, 0.0f, (float)(H - 1));
    tx0 = (uint32_t)x0 / BIN;
    tx1 = (uint32_t)x1 / BIN;
    ty0 = (uint32_t)y0 / BIN;
    ty1 = (uint32_t)y1 / BIN;
}

__global__ __launch_bounds__(256) void k_sp_tilecount(const float4 *__restrict__ sp, const uint32_t *__restrict__ vals,
                                                      const uint32_t *__restrict__ visible, uint32_t n, uint32_t W,
                                                      uint32_t H, uint32_t *__restrict__ cnt) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t c = 0;
    if (j < *visible) {
        uint32_t tx0, ty0, tx1, ty1;
        quad_tiles(sp, vals[j], W, H, tx0, ty0, tx1, ty1);
        c = (tx1 - tx0 + 1u) * (ty1 - ty0 + 1u);
    }
    cnt[j] = c;
}

// Tile pairs, balanced by pairs rather than by splats: the splats touch from one to thousands of tiles (at the
// near view a few dozen large splats hold a fifth of the pairs, next to each other at the end of the depth order).
// The pairs are cut into blocks of EMIT_PAIRS; k_pair_owner records, for each block, the splat owning its first
// pair; k_sp_emit stages the offsets and tile boxes of the block's splats in LDS and writes one pair per item,
// finding the item's splat by binary search there.  Splat j's pairs sit at off[j] .. off[j + 1] - 1, row-major
// over its tile box; the stable sort by tile then keeps each tile's splats in sorted (back-to-front) order.  The
// value is the splat itself, read directly by the blend.
constexpr uint32_t EMIT_PAIRS = 1024;
constexpr int EMIT_THREADS = 256;

__global__ __launch_bounds__(256) void k_pair_owner(const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ visible, uint32_t n,
                                                    uint32_t npairs, uint32_t *__restrict__ owner) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t vis = *visible;
    if (j >= n || j >= vis) return;
    const uint32_t s = off[j], e = j + 1u < vis ? off[j + 1u] : npairs;   // every visible splat has a pair
    for (uint64_t blk = ((uint64_t)s + EMIT_PAIRS - 1u) / EMIT_PAIRS; blk * EMIT_PAIRS < e; blk++)
        owner[blk] = j;
}

__global__ __launch_bounds__(EMIT_THREADS) void k_sp_emit(const float4 *__restrict__ sp,
                                                          const uint32_t *__restrict__ vals,
                                                          const uint32_t *__restrict__ visible, uint32_t W,
                                                          uint32_t H, const uint32_t *__restrict__ off,
                                                          const uint32_t *__restrict__ owner, uint32_t npairs,
                                                          uint32_t *__restrict__ pkeys, uint32_t *__restrict__ pvals) {
    __shared__ uint32_t s_off[EMIT_PAIRS + 1u], s_tx0[EMIT_PAIRS + 1u], s_ty0[EMIT_PAIRS + 1u],
        s_bw[EMIT_PAIRS + 1u], s_id[EMIT_PAIRS + 1u];
    const uint32_t nblocks = (uint32_t)(((uint64_t)npairs + EMIT_PAIRS - 1u) / EMIT_PAIRS);
    const uint32_t p0 = blockIdx.x * EMIT_PAIRS, p1 = min(npairs - p0, EMIT_PAIRS) + p0;
    // the block's splats: the owner of its first pair through the owner of the next block's first pair (or the
    // last visible splat); at most EMIT_PAIRS + 1, since each splat has at least one pair
    const uint32_t j0 = owner[blockIdx.x], j1 = blockIdx.x + 1u < nblocks ? owner[blockIdx.x + 1u] : *visible - 1u;
    const uint32_t ns = j1 - j0 + 1u;
    for (uint32_t k = threadIdx.x; k < ns; k += EMIT_THREADS) {
        const uint32_t id = vals[j0 + k];
        uint32_t tx0, ty0, tx1, ty1;
        quad_tiles(sp, id, W, H, tx0, ty0, tx1, ty1);
        s_off[k] = off[j0 + k];
        s_tx0[k] = tx0;
        s_ty0[k] = ty0;
        s_bw[k] = tx1 - tx0 + 1u;
        s_id[k] = id;
    }
    __syncthreads();
    const uint32_t tw = (W + BIN - 1) / BIN;
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += EMIT_THREADS) {
        uint32_t lo = 0, hi = ns - 1u;   // the last k with s_off[k] <= p (s_off[0] <= p0)
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (s_off[mid] <= p) lo = mid;
            else hi = mid - 1u;
        }
        const uint32_t t = p - s_off[lo], bw = s_bw[lo], r = t / bw;
        pkeys[p] = (s_ty0[lo] + r) * tw + s_tx0[lo] + (t - r * bw);
        pvals[p] = s_id[lo];
    }
}

__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t *__restrict__ pkeys, uint32_t npairs,
                                                     uint32_t *__restrict__ start, uint32_t *__restrict__ end) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    uint32_t k = pkeys[p];
    if (p == 0 || pkeys[p - 1] != k) start[k] = p;
    if (p + 1 == npairs || pkeys[p + 1] != k) end[k] = p + 1;
}

// back-to-front blend of a tile's splats (gps_sorted_draw.wgsl:100-124 evaluated at pixel centres)
// One workgroup per 8x8 quadrant of a 16x16 tile (four per tile, so a tile's list is spread over four CUs), four
// waves per quadrant.  The list is walked in batches of BLEND_BATCH entries.  The staging wave loads a batch's
// records while the previous batch is evaluated and keeps, in list order, only the splats whose binning box
// (quad_tiles' box, the one that decides which tiles list a splat) overlaps the quadrant, outside which a splat's
// alpha is zero as the tile lists already assume, and whose ellipse may reach a pixel centre of the quadrant
// (ellipse_touch).  Wave w then evaluates the alpha of kept splats w,
// w + 4, ... at its 64 pixels (the divisions and the exp, the costly part, independent across splats) into LDS,
// with one bit per splat saying whether any of its alphas is non-zero; wave 0 runs the back-to-front recurrence
// over the flagged splats in list order, reading them four at a time.  The arithmetic and its order per pixel
// are those of a one-thread-per-pixel walk of the whole list, so the image has the same bits.
constexpr int BLEND_QUAD = 8;                                   // quadrant side in pixels
#ifndef GPS_BLEND_WAVES
#define GPS_BLEND_WAVES 4
#endif
constexpr int BLEND_WAVES = GPS_BLEND_WAVES;
constexpr int BLEND_BATCH = 64;
constexpr int BLEND_THREADS = 64 * BLEND_WAVES;
constexpr uint32_t BLEND_STAGE = BLEND_WAVES - 1;               // the wave that loads the next batch
#ifndef GPS_BLEND_RGROUP
#define GPS_BLEND_RGROUP 4
#endif
constexpr int RGROUP = GPS_BLEND_RGROUP;                        // flagged splats per LDS round trip of the recurrence
static_assert(BLEND_QUAD * BLEND_QUAD == 64 && BIN == 2 * BLEND_QUAD, "one wave per quadrant, four per tile");
static_assert(BLEND_BATCH == 64, "a batch is staged by one wave, one flag bit per splat");

// one step of the recurrence; alpha 0 (the splat left out) keeps the colour as it is.  Computed and selected
// rather than branched on, so that consecutive steps' independent products overlap.
GPS_DEV void blend_over(float &cr, float &cg, float &cb, float alpha, float4 c) {
    const float na = 1.0f - alpha;
    const float nr = c.x * alpha + cr * na, ng = c.y * alpha + cg * na, nb = c.z * alpha + cb * na;
    const bool take = alpha > 0.0f;
    cr = take ? nr : cr;
    cg = take ? ng : cg;
    cb = take ? nb : cb;
}

GPS_DEV uint32_t take_bit(uint64_t &m, bool &has) {
    has = m != 0u;
    const uint32_t q = has ? (uint32_t)__builtin_ctzll(m) : 0u;
    m &= m - 1u;
    return q;
}

// quad_tiles' box of splat (a, b) overlaps pixel columns [qx0, qx0 + 7] and rows [qy0, qy0 + 7]; written with
// negated comparisons so that a NaN keeps the splat (its alpha is then evaluated as everywhere else)
GPS_DEV bool box_touches(float4 a, float4 b, float qx0, float qy0) {
    const float ex = a.w * (fabsf(b.x) + fabsf(b.z)) + 1.0f, ey = a.w * (fabsf(b.y) + fabsf(b.w)) + 1.0f;
    return !(floorf(a.x - ex) > qx0 + (float)(BLEND_QUAD - 1)) && !(ceilf(a.x + ex) < qx0) &&
           !(floorf(a.y - ey) > qy0 + (float)(BLEND_QUAD - 1)) && !(ceilf(a.y + ey) < qy0);
}

// The ellipse r2 <= k^2 against the quadrant's pixel centres: false only if at every centre the r2 the blend
// computes exceeds k^2 (alpha 0).  With d a centre's offset from the splat centre and, in real arithmetic,
// U = d.b1 / |b1|^2, V = d.b2 / |b2|^2, U^2 + V^2 = d^T M d is a positive semi-definite form.  Its minimum over the
// rectangle of pixel centres is bounded from below: 0 if the splat centre lies inside; otherwise the least over
// the four edges of the edge's minimum (at an end where the minimiser lies clearly outside the edge, else the
// unconstrained minimum det c^2 / M_tt, never larger), less 2^-15 T, T = (M_xx + M_yy + 2 |M_xy|) (|x| + |y|)^2
// over the rectangle: the float rounding of M (hardware reciprocals, 1 ulp) and of these sums, cancellation
// included, stays below 2^-17 T.  The blend's float u, v differ from U, V by at most
// 2^-21 (|dx| + |dy|) (1 / |b1| + 1 / |b2|) + 2^-21 (|U| + |V|); the test rejects only when the bound exceeds
// (k (1 + 2^-14) + that distance term)^2 / (1 - 2^-14), and only when the centre and both denominators lie well
// inside the float range, so that no u, v is infinite or NaN at any pixel and no term of the bound overflows.
GPS_DEV float edge_min(float c, float lo, float hi, float m_cc, float m_tt, float m_ct, float r_tt, float det) {
    const float t = -m_ct * c * r_tt, tol = 0x1p-12f * (fabsf(lo) + fabsf(hi) + 1.0f);
    if (!(t < lo - tol) && !(t > hi + tol)) return (det * r_tt) * (c * c);   // the unconstrained minimum
    const float e = t < lo ? lo : hi;
    return m_tt * e * e + 2.0f * m_ct * e * c + m_cc * c * c;
}

GPS_DEV bool ellipse_touch(float4 a, float4 b, float qx0, float qy0) {
    const float den1 = b.x * b.x + b.y * b.y, den2 = b.z * b.z + b.w * b.w;
    // in these ranges every product below stays under 2^113: no overflow, no NaN
    const bool ok = den1 > 0x1p-20f && den1 < 0x1p20f && den2 > 0x1p-20f && den2 < 0x1p20f &&
                    fabsf(a.x) < 0x1p24f && fabsf(a.y) < 0x1p24f && a.w >= 0.0f && a.w < 0x1p10f;
    if (!ok) return true;
    const float x0 = qx0 + 0.5f - a.x, x1 = qx0 + (float)BLEND_QUAD - 0.5f - a.x;
    const float y0 = qy0 + 0.5f - a.y, y1 = qy0 + (float)BLEND_QUAD - 0.5f - a.y;
    if (x0 <= 0.0f && x1 >= 0.0f && y0 <= 0.0f && y1 >= 0.0f) return true;
    const float i1 = __builtin_amdgcn_rcpf(den1 * den1), i2 = __builtin_amdgcn_rcpf(den2 * den2);
    const float mxx = b.x * b.x * i1 + b.z * b.z * i2, mxy = b.x * b.y * i1 + b.z * b.w * i2,
                myy = b.y * b.y * i1 + b.w * b.w * i2;
    if (!(mxx > 0.0f) || !(myy > 0.0f)) return true;
    const float rxx = __builtin_amdgcn_rcpf(mxx), ryy = __builtin_amdgcn_rcpf(myy);
    const float det = fmax(mxx * myy - mxy * mxy, 0.0f);
    const float f = fmin(fmin(edge_min(y0, x0, x1, myy, mxx, mxy, rxx, det), edge_min(y1, x0, x1, myy, mxx, mxy, rxx, det)),
                         fmin(edge_min(x0, y0, y1, mxx, myy, mxy, ryy, det), edge_min(x1, y0, y1, mxx, myy, mxy, ryy, det)));
    const float xm = fmax(fabsf(x0), fabsf(x1)), ym = fmax(fabsf(y0), fabsf(y1));
    const float f_lo = f - 0x1p-15f * (mxx + myy + 2.0f * fabsf(mxy)) * ((xm + ym) * (xm + ym));
    // 1 / |b| from the hardware reciprocal square root, inflated well past its error
    const float dist = 0x1p-21f * 1.01f * (xm + ym) * (__builtin_amdgcn_rsqf(den1) + __builtin_amdgcn_rsqf(den2));
    const float bound = a.w * (1.0f + 0x1p-14f) + dist;
    return !(f_lo * (1.0f - 0x1p-14f) > bound * bound);
}

// staging wave: keep the lanes whose splat touches the quadrant, packed in lane (list) order, with the splat's
// per-pixel constants: k^2 in place of k (s_a.w) and the two axis denominators (the same operations as in the
// per-pixel expressions) with their refined reciprocals, or 0 where the exact division is kept (s_d); returns the
// count
GPS_DEV uint32_t stage_batch(bool valid, float4 a, float4 b, float4 c, float qx0, float qy0, float4 *s_a,
                             float4 *s_b, float4 *s_c, float4 *s_d, uint32_t lane) {
    const bool keep = valid && box_touches(a, b, qx0, qy0) && ellipse_touch(a, b, qx0, qy0);
    const uint64_t mask = __ballot(keep);
    if (keep) {
        const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        const float den1 = b.x * b.x + b.y * b.y, den2 = b.z * b.z + b.w * b.w, kk = a.w * a.w;
        // the per-pixel divisions by den1, den2 use a reciprocal refined here where div_by gives the correctly
        // rounded quotient: denominators in [2^-60, 2^60], numerators below 2^57 (|centre| < 2^24) and not under
        // 2^-100 unless 0.  A numerator is dx b.x + dy b.y with pixel offsets 0 or above 2^-25 and the projection's
        // axis components 0 or above 1e-20 |b| >= 1e-20 2^-30, so a non-zero one stays above 2^-100; a zero one
        // gives a zero u^2 on both paths.  k^2 >= 2^-60 keeps splats at the k = 0 edge on the exact path.
        const bool fast = den1 >= 0x1p-60f && den1 <= 0x1p60f && den2 >= 0x1p-60f && den2 <= 0x1p60f &&
                          fabsf(a.x) < 0x1p24f && fabsf(a.y) < 0x1p24f && kk >= 0x1p-60f;
        s_a[r] = make_float4(a.x, a.y, a.z, kk);
        s_b[r] = b;
        s_c[r] = c;
        s_d[r] = make_float4(den1, den2, fast ? gpsm::rcp_refined(den1) : 0.0f, fast ? gpsm::rcp_refined(den2) : 0.0f);
    }
    return (uint32_t)__popcll(mask);
}

// a kept splat's r2 = u^2 + v^2 at pixel centre (qx, qy) (the two correctly rounded divisions by the staged
// denominators)
GPS_DEV float splat_r2(float4 a, float4 b, float4 d, float qx, float qy) {
    const float dx = qx - a.x, dy = qy - a.y;
    const float n1 = dx * b.x + dy * b.y, n2 = dx * b.z + dy * b.w;
    float u, v;
    if (d.z != 0.0f) {   // the staged reciprocals (the same for every lane)
        u = gpsm::div_by(n1, d.x, d.z);
        v = gpsm::div_by(n2, d.y, d.w);
    } else {
        u = n1 / d.x;
        v = n2 / d.y;
    }
    return u * u + v * v;
}

// its alpha from r2 (o the opacity, kk = k^2): 0 where it leaves the pixel unchanged (outside the image, r2 > k^2,
// or below 1/255).  k_sp_project writes k = sqrt(2 ln(o / SP_MIN_ALPHA)) with o <= 1, so here r2 <= k^2 <= 11.1 or
// r2 is NaN: exp's argument is in [-5.6, 0], where exp_mid gives exp's bits; a NaN r2 gives min(NaN, 0.99) = 0.99,
// as exp(NaN) does.
GPS_DEV float splat_alpha(float r2, float o, float kk, bool inside) {
    float alpha = 0.0f;
    if (inside && !(r2 > kk)) {
        const bool num = r2 == r2;
        const float al = num ? fmin(o * gpsm::exp_mid(num ? -0.5f * r2 : 0.0f), 0.99f) : 0.99f;
        if (!(al < SP_MIN_ALPHA)) alpha = al;
    }
    return alpha;
}

__global__ __launch_bounds__(BLEND_THREADS) void k_sp_blend(GpsFrame f, const float4 *__restrict__ sp,
                                                            const uint32_t *__restrict__ ids,
                                                            const uint32_t *__restrict__ start,
                                                            const uint32_t *__restrict__ end,
                                                            const uint32_t *__restrict__ order,
                                                            float4 *__restrict__ accum) {
    __shared__ float4 s_a[2][BLEND_BATCH], s_b[2][BLEND_BATCH], s_c[2][BLEND_BATCH], s_d[2][BLEND_BATCH];
    __shared__ float s_alpha[BLEND_BATCH][64];   // 0: the splat leaves the pixel unchanged
    __shared__ uint64_t s_any[BLEND_WAVES];      // bit q: some pixel of the quadrant takes kept splat q
    __shared__ uint32_t s_kept[2];               // kept splats per staged batch
    const uint32_t tw = (f.W + BIN - 1) / BIN;
    const uint32_t tile = order[blockIdx.x >> 2], quad = blockIdx.x & 3u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t swave = __builtin_amdgcn_readfirstlane(wave);   // the same for the wave: a scalar loop counter
    const uint32_t qx0 = (tile % tw) * BIN + (quad & 1u) * BLEND_QUAD, qy0 = (tile / tw) * BIN + (quad >> 1) * BLEND_QUAD;
    const uint32_t x = qx0 + (lane % BLEND_QUAD), y = qy0 + (lane / BLEND_QUAD);
    const float fqx0 = (float)qx0, fqy0 = (float)qy0;
    const bool inside = x < f.W && y < f.H;
    float cr = f.bg[0], cg = f.bg[1], cb = f.bg[2];
    const float qx = (float)x + 0.5f, qy = (float)y + 0.5f;
    const uint32_t s0 = start[tile], s1 = end[tile];
    uint32_t nid = 0;   // staging wave: this lane's splat in the next batch
    if (wave == BLEND_STAGE && s0 < s1) {
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a, c = a;
        const bool valid = lane < s1 - s0;
        if (valid) {
            const uint32_t id = ids[s0 + lane];
            a = sp[3u * id];
            b = sp[3u * id + 1u];
            c = sp[3u * id + 2u];
        }
        if (lane + BLEND_BATCH < s1 - s0) nid = ids[s0 + BLEND_BATCH + lane];
        const uint32_t kept = stage_batch(valid, a, b, c, fqx0, fqy0, s_a[0], s_b[0], s_c[0], s_d[0], lane);
        if (lane == 0u) s_kept[0] = kept;
    }
    __syncthreads();
    uint32_t cur = 0;
    for (uint32_t base = s0; base < s1; base += BLEND_BATCH, cur ^= 1u) {
        const uint32_t left = s1 - base, cnt = __builtin_amdgcn_readfirstlane(s_kept[cur]);
        const bool more = lane + BLEND_BATCH < left;
        float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra, rc = ra;
        uint32_t nnid = 0;
        if (wave == BLEND_STAGE) {
            if (more) {
                ra = sp[3u * nid];
                rb = sp[3u * nid + 1u];
                rc = sp[3u * nid + 2u];
            }
            if (lane + 2u * BLEND_BATCH < left) nnid = ids[base + 2u * BLEND_BATCH + lane];
        }
        uint64_t any = 0;
        // two splats per iteration: their divisions interleave; the exp is skipped where no lane needs it
        for (uint32_t q = swave; q < cnt; q += 2u * BLEND_WAVES) {
            const uint32_t q2 = q + BLEND_WAVES < cnt ? q + BLEND_WAVES : q;
            const float4 a0 = s_a[cur][q], d0 = s_d[cur][q], a1 = s_a[cur][q2], d1 = s_d[cur][q2];
            const float r20 = splat_r2(a0, s_b[cur][q], d0, qx, qy), r21 = splat_r2(a1, s_b[cur][q2], d1, qx, qy);
            const float al0 = splat_alpha(r20, a0.z, a0.w, inside);
            s_alpha[q][lane] = al0;
            if (__ballot(al0 > 0.0f)) any |= 1ull << q;
            if (q2 != q) {
                const float al1 = splat_alpha(r21, a1.z, a1.w, inside);
                s_alpha[q2][lane] = al1;
                if (__ballot(al1 > 0.0f)) any |= 1ull << q2;
            }
        }
        if (lane == 0u) s_any[wave] = any;
        __syncthreads();
        if (wave == 0u) {
            uint64_t mv = 0;
#pragma unroll
            for (int w = 0; w < BLEND_WAVES; w++) mv |= s_any[w];
            // readfirstlane returns int: widen through uint32_t, or bit 31 would fill the upper half
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(mv >> 32)),
                           lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)mv);
            uint64_t m = ((uint64_t)hi << 32) | lo;
            while (m) {
                // RGROUP flagged splats per LDS round trip; a slot past the last flagged splat (h false) reads
                // splat 0's record and takes alpha 0
                uint32_t q[RGROUP];
                bool h[RGROUP];
#pragma unroll
                for (int g = 0; g < RGROUP; g++) q[g] = take_bit(m, h[g]);
                float al[RGROUP];
                float4 c[RGROUP];
#pragma unroll
                for (int g = 0; g < RGROUP; g++) {
                    al[g] = s_alpha[q[g]][lane];
                    c[g] = s_c[cur][q[g]];
                }
#pragma unroll
                for (int g = 0; g < RGROUP; g++) blend_over(cr, cg, cb, h[g] ? al[g] : 0.0f, c[g]);
            }
        }
        if (wave == BLEND_STAGE && left > (uint32_t)BLEND_BATCH) {
            const uint32_t kept = stage_batch(more, ra, rb, rc, fqx0, fqy0, s_a[cur ^ 1u], s_b[cur ^ 1u],
                                              s_c[cur ^ 1u], s_d[cur ^ 1u], lane);
            if (lane == 0u) s_kept[cur ^ 1u] = kept;
        }
        nid = nnid;
        __syncthreads();
    }
    if (wave == 0u && inside) accum[(size_t)y * f.W + x] = make_float4(cr, cg, cb, 1.0f);
}

// The tiles in order of decreasing list length, by powers of two (the order within one power is whatever the LDS
// atomics give; each pixel is written by one workgroup either way).  The blend's workgroups take their tiles in
// this order, so the quadrants of the longest lists start first instead of behind thousands of short ones.
constexpr int ORDER_THREADS = 1024;
__global__ __launch_bounds__(ORDER_THREADS) void k_tile_order(const uint32_t *__restrict__ start,
                                                              const uint32_t *__restrict__ end, uint32_t tiles,
                                                              uint32_t *__restrict__ order) {
    __shared__ uint32_t s_cnt[33];   // slot = leading zeros of the list length: the longest lists first
    if (threadIdx.x < 33u) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < tiles; t += ORDER_THREADS) atomicAdd(&s_cnt[__clz(end[t] - start[t])], 1u);
    __syncthreads();
    if (threadIdx.x == 0u) {
        uint32_t run = 0;
        for (int b = 0; b < 33; b++) {
            const uint32_t c = s_cnt[b];
            s_cnt[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < tiles; t += ORDER_THREADS)
        order[atomicAdd(&s_cnt[__clz(end[t] - start[t])], 1u)] = t;
}

// empty tile ranges (start = end = 0): k_tile_ranges then sets those of the tiles that have pairs
__global__ void k_tile_clear(uint32_t *__restrict__ start, uint32_t *__restrict__ end, uint32_t tiles) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tiles) {
        start[i] = 0u;
        end[i] = 0u;
    }
}

unsigned nblk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ====================================================================== host
struct SortedState {
    uint32_t n = 0, W = 0, H = 0;
    float4 *sp = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr, *ticket = nullptr, *visible = nullptr;
    uint64_t *total_host = nullptr;   // the tile-pair count (inclusive sum of the per-splat tile counts), in pinned
    uint64_t *total = nullptr;        // host memory that the scan writes through its device address
    hipEvent_t scanned = nullptr;     // recorded after that scan; the host polls it
    uint64_t *status = nullptr;
    uint64_t status_cap = 0, hist_cap = 0;
    uint32_t *cnt = nullptr;          // per sorted rank: tiles touched, then exclusive offsets
    uint4 *part = nullptr;            // per projection block: probe demand (lo, hi) and visible count
    uint32_t *pk[2] = {nullptr, nullptr}, *pv[2] = {nullptr, nullptr};
    uint64_t pair_cap = 0;
    uint32_t *owner = nullptr;        // per block of EMIT_PAIRS tile pairs: the splat owning its first pair
    uint32_t *tstart = nullptr, *tend = nullptr, *torder = nullptr;
    uint32_t tiles_cap = 0;
    uint32_t epoch = 1;
    int result = 0;                   // which keys/vals buffer holds the last sort
};

SortedState *sorted_create() { return new SortedState(); }

template <typename T>
static void sfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void sorted_destroy(SortedState *s) {
    if (!s) return;
    sfree(s->sp);
    for (int k = 0; k < 2; k++) {
        sfree(s->keys[k]);
        sfree(s->vals[k]);
        sfree(s->pk[k]);
        sfree(s->pv[k]);
    }
    sfree(s->owner);
    sfree(s->hist);
    sfree(s->ticket);
    sfree(s->visible);
    if (s->total_host) (void)hipHostFree(s->total_host);
    if (s->scanned) (void)hipEventDestroy(s->scanned);
    sfree(s->status);
    sfree(s->cnt);
    sfree(s->part);
    sfree(s->tstart);
    sfree(s->tend);
    sfree(s->torder);
    delete s;
}

#define S_TRY(call)                                                                  \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return -1;                                                               \
        }                                                                            \
    } while (0)

using ProfList = std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>>;

struct SProf {
    ProfList *l;
    hipStream_t st;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    SProf(ProfList *l_, hipStream_t s_, const char *n_) : l(l_), st(s_), name(n_) {
        if (l) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, st);
        }
    }
    ~SProf() {
        if (l) {
            (void)hipEventRecord(b, st);
            l->push_back({name, {a, b}});
        }
    }
};

static int ensure_status(SortedState *s, uint64_t tiles, std::string &err) {
    if (tiles > s->status_cap) {
        sfree(s->status);
        S_TRY(hipMalloc(&s->status, tiles * 8));
        S_TRY(hipMemset(s->status, 0, tiles * 8));
        s->status_cap = tiles;
    }
    return 0;
}

// Stable LSD radix sort of n (key, value) pairs over bits [0, bits); result in buffers[result].
template <int RB>
static void radix_pass(SortedState *s, hipStream_t st, uint32_t **kb, uint32_t **vb, int cur, uint32_t n,
                       uint32_t shift, uint32_t ntiles, uint64_t hn, uint32_t stiles) {
    hipLaunchKernelGGL(k_rs_hist<RB>, dim3(ntiles), dim3(RS_BLOCK), 0, st, kb[cur], n, shift, ntiles, s->hist);
    hipLaunchKernelGGL(k_rs_scan, dim3(stiles), dim3(256), 0, st, s->hist, (uint32_t)hn, s->ticket, s->status, stiles,
                       s->epoch, (uint64_t *)nullptr);
    s->epoch = (s->epoch + 1u) & ((1u << 26) - 1u);
    if (!s->epoch) s->epoch = 1;
    hipLaunchKernelGGL(k_rs_scatter<RB>, dim3(ntiles), dim3(RS_BLOCK), 0, st, kb[cur], vb[cur], kb[cur ^ 1],
                       vb[cur ^ 1], n, shift, ntiles, s->hist);
}

// Stable LSD sort of n (key, value) pairs by the keys' bits 0..bits-1, in the fewest passes of at most 8-bit digits,
// of equal width (24 bits: three of 8; the tile pairs' 12: two of 6, a quarter of the histogram per pass).
static int radix_sort(SortedState *s, hipStream_t st, uint32_t **kb, uint32_t **vb, uint32_t n, uint32_t bits,
                      int &result, std::string &err, ProfList *prof) {
    result = 0;
    if (n <= 1 || bits == 0) return 0;
    const uint32_t npass = (bits + 7u) / 8u, rb = max((bits + npass - 1u) / npass, 4u);
    uint32_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    uint64_t hn = (1ull << rb) * ntiles;
    if (hn > s->hist_cap) {
        sfree(s->hist);
        S_TRY(hipMalloc(&s->hist, hn * 4));
        s->hist_cap = hn;
    }
    uint32_t stiles = (uint32_t)((hn + SP_SCAN_TILE - 1) / SP_SCAN_TILE);
    if (ensure_status(s, stiles, err) != 0) return -1;
    int cur = 0;
    for (uint32_t d = 0; d < npass; d++) {
        SProf p(prof, st, "radix_pass");
        const uint32_t shift = d * rb;
        if (rb == 4) radix_pass<4>(s, st, kb, vb, cur, n, shift, ntiles, hn, stiles);
        else if (rb == 5) radix_pass<5>(s, st, kb, vb, cur, n, shift, ntiles, hn, stiles);
        else if (rb == 6) radix_pass<6>(s, st, kb, vb, cur, n, shift, ntiles, hn, stiles);
        else if (rb == 7) radix_pass<7>(s, st, kb, vb, cur, n, shift, ntiles, hn, stiles);
        else radix_pass<8>(s, st, kb, vb, cur, n, shift, ntiles, hn, stiles);
        cur ^= 1;
    }
    result = cur;
    S_TRY(hipGetLastError());
    return 0;
}

int sorted_render(SortedState *s, hipStream_t st, const GpsFrame &f, const float4 *splats, const float *sh,
                  float4 *accum, uint64_t *stats, int probe_only, std::string &err, ProfList *prof) {
    const uint32_t n = f.N;
    if (n != s->n || f.W != s->W || f.H != s->H) {
        S_TRY(hipStreamSynchronize(st));
        sfree(s->sp);
        for (int k = 0; k < 2; k++) {
            sfree(s->keys[k]);
            sfree(s->vals[k]);
        }
        sfree(s->cnt);
        sfree(s->part);
        sfree(s->tstart);
        sfree(s->tend);
        sfree(s->torder);
        S_TRY(hipMalloc(&s->sp, (uint64_t)n * 48));
        for (int k = 0; k < 2; k++) {
            S_TRY(hipMalloc(&s->keys[k], (uint64_t)n * 4));
            S_TRY(hipMalloc(&s->vals[k], (uint64_t)n * 4));
        }
        S_TRY(hipMalloc(&s->cnt, (uint64_t)n * 4));
        S_TRY(hipMalloc(&s->part, (uint64_t)nblk(n, 256) * sizeof(uint4)));
        if (!s->ticket) {
            S_TRY(hipMalloc(&s->ticket, 4));
            S_TRY(hipMemset(s->ticket, 0, 4));
            S_TRY(hipMalloc(&s->visible, 4));
            S_TRY(hipHostMalloc((void **)&s->total_host, 8, hipHostMallocMapped));
            S_TRY(hipHostGetDevicePointer((void **)&s->total, s->total_host, 0));
            S_TRY(hipEventCreateWithFlags(&s->scanned, hipEventDisableTiming));
        }
        s->tiles_cap = ((f.W + BIN - 1) / BIN) * ((f.H + BIN - 1) / BIN);
        S_TRY(hipMalloc(&s->tstart, (uint64_t)s->tiles_cap * 4));
        S_TRY(hipMalloc(&s->tend, (uint64_t)s->tiles_cap * 4));
        S_TRY(hipMalloc(&s->torder, (uint64_t)s->tiles_cap * 4));
        s->n = n;
        s->W = f.W;
        s->H = f.H;
    }
    {
        SProf p(prof, st, "sp_project");
        hipLaunchKernelGGL(k_sp_project, dim3(nblk(n, 256)), dim3(256), 0, st, f, splats, sh, s->sp, s->keys[0],
                           s->vals[0], s->part, probe_only);
        hipLaunchKernelGGL(k_sp_totals, dim3(1), dim3(TOTALS_THREADS), 0, st, s->part, nblk(n, 256), stats,
                           s->visible, probe_only);
    }
    if (probe_only) {
        S_TRY(hipGetLastError());
        return 0;
    }
    int res = 0;
    if (radix_sort(s, st, s->keys, s->vals, n, 24, res, err, prof) != 0) return -1;
    s->result = res;
    uint32_t *skeys = s->keys[res], *svals = s->vals[res];
    (void)skeys;
    // binning: tiles per visible splat (sorted order) -> exclusive offsets -> pairs -> stable sort by tile
    hipLaunchKernelGGL(k_sp_tilecount, dim3(nblk(n, 256)), dim3(256), 0, st, s->sp, svals, s->visible, n, f.W, f.H,
                       s->cnt);
    uint32_t stiles = (n + SP_SCAN_TILE - 1) / SP_SCAN_TILE;
    if (ensure_status(s, stiles, err) != 0) return -1;
    hipLaunchKernelGGL(k_rs_scan, dim3(stiles), dim3(256), 0, st, s->cnt, n, s->ticket, s->status, stiles, s->epoch,
                       s->total);
    s->epoch = (s->epoch + 1u) & ((1u << 26) - 1u);
    if (!s->epoch) s->epoch = 1;
    hipLaunchKernelGGL(k_tile_clear, dim3(nblk(s->tiles_cap, 256)), dim3(256), 0, st, s->tstart, s->tend,
                       s->tiles_cap);
    // one host read per frame, the pair buffers are sized from it: poll the event (a blocking wait adds the
    // thread's wake-up to every frame) and read the count the scan wrote to pinned memory
    S_TRY(hipEventRecord(s->scanned, st));
    for (;;) {
        const hipError_t q = hipEventQuery(s->scanned);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) S_TRY(q);
    }
    const uint64_t npairs = *(volatile uint64_t *)s->total_host;
    if (npairs > 0xFFFFFFFFull) {
        err = "sorted path: too many tile pairs";
        return -1;
    }
    if (npairs > s->pair_cap) {
        for (int k = 0; k < 2; k++) {
            sfree(s->pk[k]);
            sfree(s->pv[k]);
            S_TRY(hipMalloc(&s->pk[k], npairs * 4));
            S_TRY(hipMalloc(&s->pv[k], npairs * 4));
        }
        sfree(s->owner);
        S_TRY(hipMalloc(&s->owner, ((npairs + EMIT_PAIRS - 1) / EMIT_PAIRS) * 4));
        s->pair_cap = npairs;
    }
    int pres = 0;   // which pair buffer holds the tile-sorted pairs (with none, nothing is read from it)
    if (npairs) {
        const uint32_t eblocks = nblk(npairs, EMIT_PAIRS);
        hipLaunchKernelGGL(k_pair_owner, dim3(nblk(n, 256)), dim3(256), 0, st, s->cnt, s->visible, n,
                           (uint32_t)npairs, s->owner);
        hipLaunchKernelGGL(k_sp_emit, dim3(eblocks), dim3(EMIT_THREADS), 0, st, s->sp, svals, s->visible, f.W, f.H,
                           s->cnt, s->owner, (uint32_t)npairs, s->pk[0], s->pv[0]);
        uint32_t bits = 1;   // the tile keys' width: tile ids below tiles_cap
        while (bits < 32u && (1u << bits) < s->tiles_cap) bits++;
        if (radix_sort(s, st, s->pk, s->pv, (uint32_t)npairs, bits, pres, err, prof) != 0) return -1;
        hipLaunchKernelGGL(k_tile_ranges, dim3(nblk(npairs, 256)), dim3(256), 0, st, s->pk[pres], (uint32_t)npairs,
                           s->tstart, s->tend);
    }
    {
        SProf p(prof, st, "sp_blend");
        hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(ORDER_THREADS), 0, st, s->tstart, s->tend, s->tiles_cap,
                           s->torder);
        hipLaunchKernelGGL(k_sp_blend, dim3(4u * s->tiles_cap), dim3(BLEND_THREADS), 0, st, f, s->sp,
                           s->pv[pres], s->tstart, s->tend, s->torder, accum);
    }
    S_TRY(hipGetLastError());
    return 0;
}

int sorted_read(SortedState *s, hipStream_t st, uint32_t n, uint32_t *keys, uint32_t *vals, std::string &err) {
    if (!s->keys[0] || s->n != n) {
        err = "no sorted render yet";
        return -1;
    }
    S_TRY(hipMemcpyAsync(keys, s->keys[s->result], (uint64_t)n * 4, hipMemcpyDeviceToHost, st));
    S_TRY(hipMemcpyAsync(vals, s->vals[s->result], (uint64_t)n * 4, hipMemcpyDeviceToHost, st));
    S_TRY(hipStreamSynchronize(st));
    return 0;
}
