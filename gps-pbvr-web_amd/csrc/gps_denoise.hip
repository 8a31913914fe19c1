// Learned denoiser of few-ensemble PBVR images: a 5-layer 3x3 convolution network (32 channels, ReLU,
// residual output) over 7 per-pixel features of the renderer's accumulation, BF16 operands on the gfx950
// matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, trained with Adam on pairs the renderer
// generates.  No counterpart in the reference, whose images are plain ensemble means (verify.rs:54-56).
//
// Tensors are NHWC.  Activations are bf16 with 32 channels (the feature input has 8: 7 features and a zero),
// so one 16-byte load gives a lane the 8 consecutive channels its MFMA operand needs.  Convolutions are
// implicit GEMMs: D[cout][pixel] = sum over (tap, cin) of W[cout][tap][cin] * X[pixel + tap][cin], with the
// packed weights as the A operand held in registers for the whole tile and the input tile, with its one-pixel
// halo, staged in LDS as the B operand.  The result lane then owns 4 consecutive output channels of one pixel,
// written as one 8-byte store.
//
// Kernels
//   k_dn_features   accumulation + moments + nearest depth -> 8-channel bf16 input and f32 mean
//   k_dn_conv       one convolution layer, forward (bias, ReLU) or backward-data (ReLU mask); the last layer
//                   adds the mean (residual) and, in training, writes the loss gradient and loss partials
//   k_dn_wgrad      weight and bias gradient partials per tile (f32), then k_dn_reduce sums them in tile order
//   k_dn_adam       Adam update of the f32 master weights; k_dn_pack re-packs the bf16 MFMA operands
//   k_dn_fallback   per 16x16 tile: falls back to the plain mean when the denoiser's correction is larger
//                   than the variance of the mean explains
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gps_devmath.h"
#include "gps_internal.h"
#include "../../include/gps_hip.h"

namespace {

#ifndef GPS_DN_LAYERS
#define GPS_DN_LAYERS 5
#endif
constexpr int NL = GPS_DN_LAYERS;     // layers (the first 8 -> 32 channels, the last 32 -> 3)
static_assert(NL >= 3, "at least one hidden 32 -> 32 layer");
constexpr int HID = 32;               // hidden channels
constexpr int IN_CH = 8;              // feature channels as stored (7 used)
constexpr int TW = 64, TH = 4;        // conv tile: 64 x 4 pixels, one row per wave
constexpr int HALO_W = TW + 2, HALO_H = TH + 2;
constexpr int FB_TILE = 16;           // fallback decision tile

// layer shapes: cin (as stored), cout (real)
struct LayerShapes {
    int cin[NL], cout[NL];
    constexpr LayerShapes() : cin(), cout() {
        for (int l = 0; l < NL; l++) {
            cin[l] = l == 0 ? IN_CH : HID;
            cout[l] = l == NL - 1 ? 3 : HID;
        }
    }
};
constexpr LayerShapes SHAPES{};
constexpr const int (&L_CIN)[NL] = SHAPES.cin;
constexpr const int (&L_COUT)[NL] = SHAPES.cout;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float x) {
    // round to nearest even (finite inputs)
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// K-step s, lane group g -> (tap, first channel).  cin 32: one tap per step, 8 channels per group;
// cin 8: four taps per step, one per group (tap 9..11 are zero padding).
__host__ __device__ inline void kmap(int cin, int s, int g, int &tap, int &c0) {
    if (cin == HID) {
        tap = s;
        c0 = 8 * g;
    } else {
        tap = 4 * s + g;
        c0 = 0;
    }
}
__host__ __device__ constexpr int ksteps(int cin) { return cin == HID ? 9 : 3; }

// offsets of each layer's parameters in the f32 master vector: W[cout][9][cin] then b[cout]
struct Layout {
    uint32_t w[NL], b[NL], total;
    Layout() {
        uint32_t o = 0;
        for (int l = 0; l < NL; l++) {
            w[l] = o;
            o += (uint32_t)(L_COUT[l] * 9 * L_CIN[l]);
            b[l] = o;
            o += (uint32_t)L_COUT[l];
        }
        total = o;
    }
};

// ------------------------------------------------------------------ features
// c0..2 mean rgb; c3 per-ensemble standard deviation (mean over rgb of the variance, square-rooted);
// c4 8 x standard error of the mean; c5 1/(1 + nearest depth) (0 without a hit); c6 covered fraction of
// subpixel samples; c7 zero.
// counts (4 words) are cleared for the fallback that follows.
__global__ void k_dn_features(const float4 *__restrict__ accum, const float4 *__restrict__ mom,
                              const float *__restrict__ zmin, uint32_t X, float spp, uint16_t *__restrict__ x8,
                              float4 *__restrict__ mean, uint32_t *__restrict__ counts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4u) counts[i] = 0u;
    if (i >= X) return;
    const float4 a = accum[i];
    const float4 m = mom[i];
    const float n = a.w > 1.0f ? a.w : 1.0f;
    const float mr = a.x / n, mg = a.y / n, mb = a.z / n;
    const float vr = fmaxf(m.x / n - mr * mr, 0.0f);
    const float vg = fmaxf(m.y / n - mg * mg, 0.0f);
    const float vb = fmaxf(m.z / n - mb * mb, 0.0f);
    const float var = (vr + vg + vb) / 3.0f;
    const float z = zmin[i];
    float f[8];
    f[0] = mr;
    f[1] = mg;
    f[2] = mb;
    f[3] = sqrtf(var);
    f[4] = 8.0f * sqrtf(var / n);
    f[5] = z < 3.0e38f ? 1.0f / (1.0f + z) : 0.0f;
    f[6] = m.w / (n * spp);
    f[7] = 0.0f;
    uint4 o;
    o.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
    o.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
    o.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
    o.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
    reinterpret_cast<uint4 *>(x8)[i] = o;
    // variance of the mean per channel in w (for the fallback test)
    mean[i] = make_float4(mr, mg, mb, (vr + vg + vb) / n);
}

// ------------------------------------------------------------------ convolution (forward / backward-data)
enum { EPI_RELU = 0, EPI_MASK = 1, EPI_RESID = 2 };

// Grid: (ceil(W / 64), ceil(H / 4), images).  Block: 256 threads, wave w computes tile row w.
// in:  [img][H][W][CIN] bf16; wpk: packed A operands [MT][KS][64 lanes][8] bf16; bias: [cout] f32 or null
// EPI_RELU:  out [img][H][W][32] bf16 = relu(conv + bias)
// EPI_MASK:  out [img][H][W][32] bf16 = conv * (mask > 0)     (mask: forward input of the layer)
// EPI_RESID: out4 [img][H][W] float4 = (mean.rgb + conv + bias, 1); with target: dy [img][H][W][8] bf16 the
//            loss gradient 2 (out - target) * gscale * iweight[img] and the weighted loss partial sums per
//            block in lpart
template <int CIN, int MT, int EPI>
__global__ __launch_bounds__(256) void k_dn_conv(const uint16_t *__restrict__ in, uint32_t W, uint32_t H,
                                                 const uint4 *__restrict__ wpk, const float *__restrict__ bias,
                                                 uint16_t *__restrict__ out, const uint16_t *__restrict__ mask,
                                                 const float4 *__restrict__ mean, float4 *__restrict__ out4,
                                                 const float4 *__restrict__ target, uint16_t *__restrict__ dy,
                                                 float gscale, float *__restrict__ lpart,
                                                 const float *__restrict__ iweight) {
    constexpr int KS = ksteps(CIN);
    constexpr int PSTRIDE = CIN == HID ? 40 : 8;  // LDS pixel stride in bf16 (80 B: conflict-free 16 B reads)
    __shared__ __align__(16) uint16_t tile[HALO_H * HALO_W * PSTRIDE];
    __shared__ float red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const uint32_t img = blockIdx.z;
    const int x0 = (int)blockIdx.x * TW, y0 = (int)blockIdx.y * TH;
    const uint16_t *src = in + (size_t)img * H * W * CIN;

    // A operands (weights) for this lane, held in registers
    bf16x8 A[MT][KS];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int s = 0; s < KS; s++) {
            uint4 v = wpk[((size_t)mt * KS + s) * 64 + lane];
            A[mt][s] = *reinterpret_cast<bf16x8 *>(&v);
        }

    // stage the input tile with its halo (zero outside the image)
    constexpr int CHUNKS = CIN / 8;  // 16-byte chunks per pixel
    for (int t = threadIdx.x; t < HALO_H * HALO_W * CHUNKS; t += 256) {
        const int c = t % CHUNKS, p = t / CHUNKS;
        const int hx = p % HALO_W, hy = p / HALO_W;
        const int gx = x0 + hx - 1, gy = y0 + hy - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gx >= 0 && gy >= 0 && gx < (int)W && gy < (int)H)
            v = reinterpret_cast<const uint4 *>(src + ((size_t)gy * W + gx) * CIN)[c];
        *reinterpret_cast<uint4 *>(&tile[p * PSTRIDE + c * 8]) = v;
    }
    __syncthreads();

    f32x4 acc[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < KS; s++) {
        int tap, c0;
        kmap(CIN, s, g, tap, c0);
        const int dy0 = tap / 3, dx0 = tap % 3;
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            bf16x8 B;
            if (tap < 9) {
                const int p = (wave + dy0) * HALO_W + nt * 16 + n + dx0;
                uint4 v = *reinterpret_cast<const uint4 *>(&tile[p * PSTRIDE + c0]);
                B = *reinterpret_cast<bf16x8 *>(&v);
            } else {
                uint4 z = make_uint4(0u, 0u, 0u, 0u);
                B = *reinterpret_cast<bf16x8 *>(&z);
            }
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt][s], B, acc[mt][nt], 0, 0, 0);
        }
    }

    // epilogue: lane holds output channels mt*16 + 4g + v of pixel nt*16 + n of row `wave`
    const int y = y0 + wave;
    float lsum = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const int x = x0 + nt * 16 + n;
        if (x >= (int)W || y >= (int)H) continue;
        const size_t pix = ((size_t)img * H + y) * W + x;
        if (EPI == EPI_RESID) {
            if (g != 0) continue;  // channels 0..2 live in lane group 0
            const float4 m = mean[pix];
            float o0 = m.x + (acc[0][nt][0] + bias[0]);
            float o1 = m.y + (acc[0][nt][1] + bias[1]);
            float o2 = m.z + (acc[0][nt][2] + bias[2]);
            out4[pix] = make_float4(o0, o1, o2, 1.0f);
            if (target) {
                const float4 t = target[pix];
                const float e0 = o0 - t.x, e1 = o1 - t.y, e2 = o2 - t.z;
                const float wi = iweight[img];
                lsum += wi * (e0 * e0 + e1 * e1 + e2 * e2);
                const float gs = 2.0f * gscale * wi;
                uint4 d;
                d.x = (uint32_t)f2bf(e0 * gs) | ((uint32_t)f2bf(e1 * gs) << 16);
                d.y = (uint32_t)f2bf(e2 * gs);
                d.z = 0u;
                d.w = 0u;
                reinterpret_cast<uint4 *>(dy)[pix] = d;
            }
        } else {
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                const int c = mt * 16 + 4 * g;
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = acc[mt][nt][k];
                if (EPI == EPI_RELU) {
#pragma unroll
                    for (int k = 0; k < 4; k++) v[k] = fmaxf(v[k] + bias[c + k], 0.0f);
                } else {
                    const uint2 mk = *reinterpret_cast<const uint2 *>(mask + pix * HID + c);
                    const uint32_t mw[4] = {mk.x & 0xFFFFu, mk.x >> 16, mk.y & 0xFFFFu, mk.y >> 16};
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (!(bf2f((uint16_t)mw[k]) > 0.0f)) v[k] = 0.0f;
                }
                uint2 o;
                if (EPI == EPI_RELU) {   // forward activations: the pack the fused forward uses
                    o.x = gpsm::bf16_pack2(v[0], v[1]);
                    o.y = gpsm::bf16_pack2(v[2], v[3]);
                } else {
                    o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                    o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                }
                *reinterpret_cast<uint2 *>(out + pix * HID + c) = o;
            }
        }
    }
    if (EPI == EPI_RESID && target) {
        // block sum of the squared errors, fixed order
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (lane == 0) red[wave] = lsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
            lpart[b] = (red[0] + red[1]) + (red[2] + red[3]);
        }
    }
}

// ------------------------------------------------------------------ fused forward (inference)
// All NL layers of one FW x FH output tile in one workgroup.  The feature tile with its NL-pixel halo is staged
// once; each layer's output over the tile, grown by the halo the later layers still need, stays in LDS as bf16
// (the values the layer-by-layer kernels store, so both forwards give the same bits).  Layer l's output region is
// (FW + 2 (NL - 1 - l)) x (FH + 2 (NL - 1 - l)) pixels; its pixels outside the image are stored as zero, the
// padding the next layer's taps must see.  A region's pixels in row-major order form the 16-pixel columns of the
// MFMA B operands; wave w computes columns w, w + FWAVES, ..., each column's nine B operands loaded before its
// MFMAs.  Two activation buffers alternate: FA holds the outputs of the even layers, FB the features and the outputs
// of the odd layers.  A layer's packed weights and biases are read from one of two LDS
// buffers; the next layer's are loaded from memory while the layer computes and stored to the other buffer
// before the layer's closing barrier.  One workgroup per CU (the LDS), sixteen waves.
#ifndef GPS_DN_FWAVES
#define GPS_DN_FWAVES 16   // at 1280 x 720: 16 waves 0.155 ms per forward, 12 waves 0.160, 8 waves 0.182 (measured
                           // before the weight prefetch moved out of scratch; 16 waves now 0.141)
#endif
constexpr int FW = 32, FH = 16, FWAVES = GPS_DN_FWAVES, FTHREADS = FWAVES * 64;
constexpr int fr_w(int l) { return FW + 2 * (NL - 1 - l); }
constexpr int fr_h(int l) { return FH + 2 * (NL - 1 - l); }
constexpr int FIN_W = FW + 2 * NL, FIN_H = FH + 2 * NL;
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// 32-channel regions: 64 B per pixel, its four 16 B chunks (channels 8k..8k+7) at chunk positions (k ^ (p >> 1)) & 3,
// so the 16-pixel, one-chunk reads of a B operand meet no LDS bank conflict (an unswizzled 80 B stride conflicted on
// half of the LDS cycles on gfx950)
__device__ __forceinline__ int fsw(int p, int chunk) { return p * HID + (((chunk ^ (p >> 1)) & 3) << 3); }
constexpr int F_A = fr_w(0) * fr_h(0) * HID;
constexpr int F_B = cmax(FIN_W * FIN_H * IN_CH, fr_w(1) * fr_h(1) * HID);
constexpr int fwt_a(int l) { return ((l == NL - 1 ? 3 : HID) + 15) / 16 * ksteps(l == 0 ? IN_CH : HID) * 64; }
constexpr int FWT = 2 * 9 * 64 + HID / 4;   // uint4 per weight buffer: a 32 -> 32 layer's A operands, then 32 biases
constexpr int FWT_LOADS = (FWT + FTHREADS - 1) / FTHREADS;
constexpr bool FUSED_FITS = (F_A + F_B) * 2 + 2 * FWT * 16 <= 160 * 1024;   // deeper networks: layer by layer

// per layer: the packed A operands followed by the biases (HID floats, zero past cout), as pack_all writes them
struct FusedNet {
    const uint4 *w[NL];
};

template <int L>
__device__ __forceinline__ void fused_fetch(const FusedNet &net, uint4 (&r)[FWT_LOADS]) {
    if constexpr (L < NL) {
        constexpr int N = fwt_a(L) + HID / 4;
#pragma unroll
        for (int i = 0; i < FWT_LOADS; i++) {
            const int t = threadIdx.x + i * FTHREADS;
            r[i] = net.w[L][t < N ? t : N - 1];   // every element written: the array stays in registers
        }
    }
}
template <int L>
__device__ __forceinline__ void fused_store(uint4 *__restrict__ sw, const uint4 (&r)[FWT_LOADS]) {
    if constexpr (L < NL) {
        constexpr int N = fwt_a(L) + HID / 4;
#pragma unroll
        for (int i = 0; i < FWT_LOADS; i++) {
            const int t = threadIdx.x + i * FTHREADS;
            if (t < N) sw[t] = r[i];
        }
    }
}

template <int L>
__device__ __forceinline__ void fused_layer(const uint16_t *__restrict__ src, uint16_t *__restrict__ dst,
                                            const uint4 *__restrict__ wl, int x0, int y0, uint32_t W, uint32_t H,
                                            const float4 *__restrict__ mean, float4 *__restrict__ out4) {
    constexpr bool LAST = L == NL - 1;
    constexpr int CIN = L == 0 ? IN_CH : HID, MT = LAST ? 1 : 2;
    constexpr int RW = fr_w(L), R = RW * fr_h(L), SW = RW + 2;   // output region; its input is SW wide
    constexpr int KS = ksteps(CIN);
    constexpr int NT = (R + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int rx0 = x0 - (NL - 1 - L), ry0 = y0 - (NL - 1 - L);   // image position of the region's first pixel
    bf16x8 A[MT][KS];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int s = 0; s < KS; s++) {
            uint4 v = wl[(mt * KS + s) * 64 + lane];
            A[mt][s] = *reinterpret_cast<bf16x8 *>(&v);
        }
    const float *bl = reinterpret_cast<const float *>(wl + MT * KS * 64);
    float bias[MT][4];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int k = 0; k < 4; k++) bias[mt][k] = bl[LAST ? k : mt * 16 + 4 * g + k];
    for (int nt = wave; nt < NT; nt += FWAVES) {
        const int q = nt * 16 + n;
        const int qq = q < R ? q : R - 1;   // lanes past the region compute a copy, never stored
        const int oy = qq / RW, ox = qq - oy * RW;
        const int base = oy * SW + ox;      // input pixel under tap (0, 0)
        bf16x8 B[KS];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            int tap, c0;
            kmap(CIN, s, g, tap, c0);
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (tap < 9) {
                const int p = base + (tap / 3) * SW + tap % 3;
                v = *reinterpret_cast<const uint4 *>(&src[CIN == HID ? fsw(p, g) : p * IN_CH]);
            }
            B[s] = *reinterpret_cast<bf16x8 *>(&v);
        }
        f32x4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < KS; s++)
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt][s], B[s], acc[mt], 0, 0, 0);
        // epilogue: lane holds output channels mt*16 + 4g + k of region pixel q
        const int gx = rx0 + ox, gy = ry0 + oy;
        const bool inside = (uint32_t)gx < W && (uint32_t)gy < H;
        if (LAST) {
            if (q < R && g == 0 && inside) {   // channels 0..2 live in lane group 0
                const size_t pix = (size_t)gy * W + gx;
                const float4 m = mean[pix];
                out4[pix] = make_float4(m.x + (acc[0][0] + bias[0][0]), m.y + (acc[0][1] + bias[0][1]),
                                        m.z + (acc[0][2] + bias[0][2]), 1.0f);
            }
        } else if (q < R) {
#pragma unroll
            for (int mt = 0; mt < MT; mt++) {
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = inside ? fmaxf(acc[mt][k] + bias[mt][k], 0.0f) : 0.0f;
                const uint2 o = make_uint2(gpsm::bf16_pack2(v[0], v[1]), gpsm::bf16_pack2(v[2], v[3]));
                *reinterpret_cast<uint2 *>(&dst[fsw(q, 2 * mt + (g >> 1)) + 4 * (g & 1)]) = o;
            }
        }
    }
}

// layers L..NL-1: even layers read FB and write FA, odd layers the reverse; the last writes the image.  Layer L's
// weights are in weight buffer L % 2; layer L + 1's are fetched during layer L into the other one.
template <int L>
__device__ __forceinline__ void fused_layers(uint16_t *sa, uint16_t *sb, uint4 *sw, const FusedNet &net, int x0,
                                             int y0, uint32_t W, uint32_t H, const float4 *__restrict__ mean,
                                             float4 *__restrict__ out4) {
    if constexpr (L < NL) {
        uint4 nxt[FWT_LOADS];
        fused_fetch<L + 1>(net, nxt);
        fused_layer<L>(L % 2 == 0 ? sb : sa, L % 2 == 0 ? sa : sb, sw + (L % 2) * FWT, x0, y0, W, H, mean, out4);
        if constexpr (L < NL - 1) {
            fused_store<L + 1>(sw + ((L + 1) % 2) * FWT, nxt);
            __syncthreads();
            fused_layers<L + 1>(sa, sb, sw, net, x0, y0, W, H, mean, out4);
        }
    }
}

// Grid: (ceil(W / FW), ceil(H / FH)).  in: [H][W][8] bf16 features; mean: [H][W] float4; out4: [H][W] float4.
__global__ __launch_bounds__(FTHREADS) void k_dn_fused(const uint16_t *__restrict__ in, uint32_t W, uint32_t H,
                                                       FusedNet net, const float4 *__restrict__ mean,
                                                       float4 *__restrict__ out4) {
    __shared__ __align__(16) uint16_t sa[FUSED_FITS ? F_A : 8];
    __shared__ __align__(16) uint16_t sb[FUSED_FITS ? F_B : 8];
    __shared__ __align__(16) uint4 sw[FUSED_FITS ? 2 * FWT : 1];
    if constexpr (FUSED_FITS) {
        const int x0 = (int)blockIdx.x * FW, y0 = (int)blockIdx.y * FH;
        uint4 w0[FWT_LOADS];
        fused_fetch<0>(net, w0);
        for (int t = threadIdx.x; t < FIN_W * FIN_H; t += FTHREADS) {
            const int gx = x0 - NL + t % FIN_W, gy = y0 - NL + t / FIN_W;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gx >= 0 && gy >= 0 && gx < (int)W && gy < (int)H)
                v = reinterpret_cast<const uint4 *>(in)[(size_t)gy * W + gx];
            *reinterpret_cast<uint4 *>(&sb[t * IN_CH]) = v;
        }
        fused_store<0>(sw, w0);
        __syncthreads();
        fused_layers<0>(sa, sb, sw, net, x0, y0, W, H, mean, out4);
    }
}

// ------------------------------------------------------------------ weight gradient
// Per tile of 64 x 4 pixels: dW[cout][tap][cin] = sum over the tile's pixels of dy[p][cout] * x[p + tap][cin]
// and db[cout] = sum of dy[p][cout], written to part[block][...] (layout of the master weights of the layer).
// Thread t: cout = t / 8, input channels (t % 8) * CIN / 8 + j.
template <int CIN, int COUT, int DYC>
__global__ __launch_bounds__(256) void k_dn_wgrad(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dy,
                                                  uint32_t W, uint32_t H, float *__restrict__ part) {
    constexpr int CPT = CIN / 8;  // input channels per thread
    __shared__ uint16_t xs[HALO_H * HALO_W][CIN + 2];
    __shared__ uint16_t ds[TH * TW][COUT + 1];
    const uint32_t img = blockIdx.z;
    const int x0 = (int)blockIdx.x * TW, y0 = (int)blockIdx.y * TH;
    const uint16_t *xsrc = x + (size_t)img * H * W * CIN;
    const uint16_t *dsrc = dy + (size_t)img * H * W * DYC;
    for (int t = threadIdx.x; t < HALO_H * HALO_W * CIN; t += 256) {
        const int c = t % CIN, p = t / CIN;
        const int gx = x0 + p % HALO_W - 1, gy = y0 + p / HALO_W - 1;
        uint16_t v = 0;
        if (gx >= 0 && gy >= 0 && gx < (int)W && gy < (int)H) v = xsrc[((size_t)gy * W + gx) * CIN + c];
        xs[p][c] = v;
    }
    for (int t = threadIdx.x; t < TH * TW * COUT; t += 256) {
        const int c = t % COUT, p = t / COUT;
        const int gx = x0 + p % TW, gy = y0 + p / TW;
        uint16_t v = 0;
        if (gx < (int)W && gy < (int)H) v = dsrc[((size_t)gy * W + gx) * DYC + c];
        ds[p][c] = v;
    }
    __syncthreads();
    const int co = threadIdx.x / 8, cb = (threadIdx.x % 8) * CPT;
    const uint32_t blk = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    constexpr int WSZ = COUT * 9 * CIN + COUT;
    float *dst = part + (size_t)blk * WSZ;
    if (co < COUT) {
        float a[9][CPT];
        for (int t = 0; t < 9; t++)
            for (int j = 0; j < CPT; j++) a[t][j] = 0.0f;
        float bsum = 0.0f;
        for (int p = 0; p < TH * TW; p++) {
            const float d = bf2f(ds[p][co]);
            bsum += d;
            const int px = p % TW, py = p / TW;
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int hp = (py + t / 3) * HALO_W + px + t % 3;
#pragma unroll
                for (int j = 0; j < CPT; j++) a[t][j] += d * bf2f(xs[hp][cb + j]);
            }
        }
        for (int t = 0; t < 9; t++)
            for (int j = 0; j < CPT; j++) dst[((size_t)co * 9 + t) * CIN + cb + j] = a[t][j];
        if ((threadIdx.x % 8) == 0) dst[COUT * 9 * CIN + co] = bsum;
    }
}

// sums `parts` partial vectors of length n (fixed order) into out
__global__ void k_dn_reduce(const float *__restrict__ part, uint32_t parts, uint32_t n, float *__restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t b = 0; b < parts; b++) s += part[(size_t)b * n + i];
    out[i] = s;
}

__global__ void k_dn_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                          float c2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Packs one layer's A operands.  Forward: rows = cout, K = (tap, cin).  Backward-data: rows = cin of the
// layer, K = (tap', cout) with tap' = 8 - tap (the transposed, flipped kernel); its stored input channel count
// is 8 for the last layer (3 real) and 32 otherwise.
__global__ void k_dn_pack(const float *__restrict__ w, int cin, int cout, int backward, int rows, int kcin,
                          uint4 *__restrict__ out) {
    const int MT = (rows + 15) / 16, KS = ksteps(kcin);
    const int total = MT * KS * 64;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int lane = t % 64, s = (t / 64) % KS, mt = t / (64 * KS);
    const int row = mt * 16 + (lane & 15), g = lane >> 4;
    int tap, c0;
    kmap(kcin, s, g, tap, c0);
    uint16_t h[8];
    for (int j = 0; j < 8; j++) {
        float val = 0.0f;
        const int kc = c0 + j;
        if (tap < 9 && row < rows) {
            if (!backward) {
                if (kc < cin) val = w[((size_t)row * 9 + tap) * cin + kc];
            } else {
                // row = input channel of the layer, kc = output channel of the layer
                if (kc < cout && row < cin) val = w[((size_t)kc * 9 + (8 - tap)) * cin + row];
            }
        }
        h[j] = f2bf(val);
    }
    uint4 o;
    o.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    o.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    o.z = (uint32_t)h[4] | ((uint32_t)h[5] << 16);
    o.w = (uint32_t)h[6] | ((uint32_t)h[7] << 16);
    out[t] = o;
}

// the biases of a layer after its packed forward operands (HID floats, zero past cout)
__global__ void k_dn_bias_tail(const float *__restrict__ b, int cout, float *__restrict__ out) {
    if (threadIdx.x < HID) out[threadIdx.x] = (int)threadIdx.x < cout ? b[threadIdx.x] : 0.0f;
}

// Per 16x16 tile: ratio = mean over pixels and channels of (denoised - mean)^2 over the mean of the variance
// of the mean.  Above `limit` the tile takes the plain mean.  counts[0] += 1 per fallback tile.
__global__ void k_dn_fallback(float4 *__restrict__ out, const float4 *__restrict__ mean, uint32_t W, uint32_t H,
                              float limit, uint32_t *__restrict__ counts, float *__restrict__ ratios) {
    __shared__ float s_num[256], s_den[256];
    const int tx = threadIdx.x % FB_TILE, ty = threadIdx.x / FB_TILE;
    const int x = blockIdx.x * FB_TILE + tx, y = blockIdx.y * FB_TILE + ty;
    float num = 0.0f, den = 0.0f;
    const bool in = x < (int)W && y < (int)H;
    size_t i = in ? (size_t)y * W + x : 0;
    if (in) {
        const float4 o = out[i], m = mean[i];
        const float a = o.x - m.x, b = o.y - m.y, c = o.z - m.z;
        num = (a * a + b * b + c * c) / 3.0f;
        den = m.w / 3.0f;
    }
    s_num[threadIdx.x] = num;
    s_den[threadIdx.x] = den;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s_num[threadIdx.x] += s_num[threadIdx.x + off];
            s_den[threadIdx.x] += s_den[threadIdx.x + off];
        }
        __syncthreads();
    }
    const float ratio = s_num[0] / fmaxf(s_den[0], 1e-30f);
    const bool fb = ratio > limit;
    if (threadIdx.x == 0) {
        ratios[blockIdx.y * gridDim.x + blockIdx.x] = ratio;
        if (fb) atomicAdd(counts, 1u);
    }
    if (fb && in) {
        const float4 m = mean[i];
        out[i] = make_float4(m.x, m.y, m.z, 1.0f);
    }
}

// gathers training pairs into a batch: per image, X (8 ch bf16), mean, target
__global__ void k_dn_gather(const uint16_t *__restrict__ px, const float4 *__restrict__ pmean,
                            const float4 *__restrict__ ptgt, const float *__restrict__ pweight,
                            const uint32_t *__restrict__ idx, uint32_t X, uint16_t *__restrict__ bx,
                            float4 *__restrict__ bmean, float4 *__restrict__ btgt, float *__restrict__ bweight) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
    if (i == 0) bweight[b] = pweight[idx[b]];
    if (i >= X) return;
    const size_t src = (size_t)idx[b] * X + i, dst = (size_t)b * X + i;
    reinterpret_cast<uint4 *>(bx)[dst] = reinterpret_cast<const uint4 *>(px)[src];
    bmean[dst] = pmean[src];
    btgt[dst] = ptgt[src];
}

__global__ void k_dn_mean_of(const float4 *__restrict__ accum, uint32_t X, float4 *__restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= X) return;
    const float4 a = accum[i];
    const float n = a.w > 1.0f ? a.w : 1.0f;
    out[i] = make_float4(a.x / n, a.y / n, a.z / n, 1.0f);
}

__global__ void k_dn_sum(const float *__restrict__ part, uint32_t n, float scale, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.0f;
        for (uint32_t i = 0; i < n; i++) s += part[i];
        *out = s * scale;
    }
}

inline unsigned cdiv(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

// host PCG32 (O'Neill) for initialisation and batch sampling
struct Pcg {
    uint64_t s, inc;
    explicit Pcg(uint64_t seed, uint64_t seq = 54u) : s(0), inc((seq << 1u) | 1u) {
        next();
        s += seed;
        next();
    }
    uint32_t next() {
        uint64_t o = s;
        s = o * 6364136223846793005ULL + inc;
        uint32_t x = (uint32_t)(((o >> 18u) ^ o) >> 27u), r = (uint32_t)(o >> 59u);
        return (x >> r) | (x << ((-r) & 31));
    }
    double uniform() { return (next() + 0.5) * (1.0 / 4294967296.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

}  // namespace

// ====================================================================== host object
struct GpsDenoiser {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Layout lay;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // master weights, Adam moments, gradient
    uint4 *fwd[NL] = {}, *bwd[NL] = {};                              // packed bf16 operands
    uint32_t adam_t = 0;
    // activations: a[0] = input (8 ch), a[l] = output of layer l-1 (32 ch); gradient ping-pong
    uint32_t cap_px = 0;  // pixels (images x W x H) the buffers hold
    uint16_t *act[NL] = {};
    uint16_t *dy8 = nullptr, *gA = nullptr, *gB = nullptr;
    float4 *mean = nullptr, *out = nullptr, *tgt = nullptr;
    float *part = nullptr, *lpart = nullptr, *ratios = nullptr;
    uint64_t part_cap = 0, lpart_cap = 0, ratio_cap = 0;
    uint32_t *counts = nullptr;
    uint32_t last_w = 0, last_h = 0;
    bool fused = true;  // inference forward: one fused kernel (FUSED_FITS) or the layer-by-layer kernels
    // training pairs
    uint32_t pair_w = 0, pair_h = 0, pairs = 0, pair_cap = 0;
    uint16_t *px = nullptr;
    float4 *pmean = nullptr, *ptgt = nullptr;
    float *pweight = nullptr, *bweight = nullptr;
    uint32_t *bidx = nullptr;
    uint32_t bidx_cap = 0;
};

namespace {

int dfail(GpsDenoiser *d, const std::string &m) {
    d->err = m;
    return -1;
}
#define DN_TRY(d, x)                                                                   \
    do {                                                                               \
        hipError_t _e = (x);                                                           \
        if (_e != hipSuccess) return dfail(d, std::string(#x) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <class T>
void dnfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

int pack_all(GpsDenoiser *d) {
    for (int l = 0; l < NL; l++) {
        const int cin = L_CIN[l], cout = L_COUT[l];
        // forward: rows = cout (padded to 16), K over the stored cin
        int rows = cout, kcin = cin;
        int total = ((rows + 15) / 16) * ksteps(kcin) * 64;
        hipLaunchKernelGGL(k_dn_pack, dim3(cdiv(total, 256)), dim3(256), 0, d->stream, d->w + d->lay.w[l], cin, cout,
                           0, rows, kcin, d->fwd[l]);
        hipLaunchKernelGGL(k_dn_bias_tail, dim3(1), dim3(HID), 0, d->stream, d->w + d->lay.b[l], cout,
                           reinterpret_cast<float *>(d->fwd[l] + total));
        if (l > 0) {
            rows = cin;
            kcin = cout == 3 ? IN_CH : HID;
            total = ((rows + 15) / 16) * ksteps(kcin) * 64;
            hipLaunchKernelGGL(k_dn_pack, dim3(cdiv(total, 256)), dim3(256), 0, d->stream, d->w + d->lay.w[l], cin,
                               cout, 1, rows, kcin, d->bwd[l]);
        }
    }
    DN_TRY(d, hipGetLastError());
    return 0;
}

int ensure_act(GpsDenoiser *d, uint32_t px, uint32_t imgs, uint32_t W, uint32_t H) {
    if (px > d->cap_px) {
        for (int l = 0; l < NL; l++) dnfree(d->act[l]);
        dnfree(d->dy8);
        dnfree(d->gA);
        dnfree(d->gB);
        dnfree(d->mean);
        dnfree(d->out);
        dnfree(d->tgt);
        DN_TRY(d, hipMalloc(&d->act[0], (size_t)px * IN_CH * 2));
        for (int l = 1; l < NL; l++) DN_TRY(d, hipMalloc(&d->act[l], (size_t)px * HID * 2));
        DN_TRY(d, hipMalloc(&d->dy8, (size_t)px * IN_CH * 2));
        DN_TRY(d, hipMalloc(&d->gA, (size_t)px * HID * 2));
        DN_TRY(d, hipMalloc(&d->gB, (size_t)px * HID * 2));
        DN_TRY(d, hipMalloc(&d->mean, (size_t)px * sizeof(float4)));
        DN_TRY(d, hipMalloc(&d->out, (size_t)px * sizeof(float4)));
        DN_TRY(d, hipMalloc(&d->tgt, (size_t)px * sizeof(float4)));
        d->cap_px = px;
    }
    const uint64_t blocks = (uint64_t)cdiv(W, TW) * cdiv(H, TH) * imgs;
    const uint64_t pneed = blocks * (uint64_t)(HID * 9 * HID + HID);
    if (pneed > d->part_cap) {
        dnfree(d->part);
        DN_TRY(d, hipMalloc(&d->part, pneed * sizeof(float)));
        d->part_cap = pneed;
    }
    if (blocks > d->lpart_cap) {
        dnfree(d->lpart);
        DN_TRY(d, hipMalloc(&d->lpart, blocks * sizeof(float)));
        d->lpart_cap = blocks;
    }
    const uint64_t tiles = (uint64_t)cdiv(W, FB_TILE) * cdiv(H, FB_TILE);
    if (tiles > d->ratio_cap) {
        dnfree(d->ratios);
        DN_TRY(d, hipMalloc(&d->ratios, tiles * sizeof(float)));
        d->ratio_cap = tiles;
    }
    return 0;
}

template <int CIN, int MT, int EPI>
void conv(GpsDenoiser *d, const uint16_t *in, uint32_t W, uint32_t H, uint32_t imgs, const uint4 *wpk,
          const float *bias, uint16_t *out, const uint16_t *mask, const float4 *mean, float4 *out4,
          const float4 *target, uint16_t *dy, float gscale, float *lpart, const float *iweight = nullptr) {
    dim3 grid(cdiv(W, TW), cdiv(H, TH), imgs);
    hipLaunchKernelGGL((k_dn_conv<CIN, MT, EPI>), grid, dim3(256), 0, d->stream, in, W, H, wpk, bias, out, mask,
                       mean, out4, target, dy, gscale, lpart, iweight);
}

// forward over `imgs` images of W x H in d->act[0] / d->mean; output in d->out.  With target: loss gradient
// in d->dy8 and loss partials in d->lpart.
void forward(GpsDenoiser *d, uint32_t W, uint32_t H, uint32_t imgs, const float4 *target, float gscale) {
    const Layout &L = d->lay;
    conv<IN_CH, 2, EPI_RELU>(d, d->act[0], W, H, imgs, d->fwd[0], d->w + L.b[0], d->act[1], nullptr, nullptr,
                             nullptr, nullptr, nullptr, 0.0f, nullptr);
    for (int l = 1; l < NL - 1; l++)
        conv<HID, 2, EPI_RELU>(d, d->act[l], W, H, imgs, d->fwd[l], d->w + L.b[l], d->act[l + 1], nullptr,
                               nullptr, nullptr, nullptr, nullptr, 0.0f, nullptr);
    conv<HID, 1, EPI_RESID>(d, d->act[NL - 1], W, H, imgs, d->fwd[NL - 1], d->w + L.b[NL - 1], nullptr, nullptr,
                            d->mean, d->out, target, d->dy8, gscale, d->lpart, d->bweight);
}

// inference forward of one W x H image in d->act[0] / d->mean into d->out: the fused kernel when the network's
// regions fit the LDS and the handle asks for it, else the layer-by-layer kernels (the same bits)
void forward_infer(GpsDenoiser *d, uint32_t W, uint32_t H) {
    if (FUSED_FITS && d->fused) {
        FusedNet net;
        for (int l = 0; l < NL; l++) net.w[l] = d->fwd[l];
        hipLaunchKernelGGL(k_dn_fused, dim3(cdiv(W, FW), cdiv(H, FH)), dim3(FTHREADS), 0, d->stream, d->act[0], W, H,
                           net, d->mean, d->out);
    } else {
        forward(d, W, H, 1, nullptr, 0.0f);
    }
}

// backward from d->dy8 (loss gradient of the output) to the gradient of every parameter in d->g
void backward(GpsDenoiser *d, uint32_t W, uint32_t H, uint32_t imgs) {
    const Layout &L = d->lay;
    dim3 grid(cdiv(W, TW), cdiv(H, TH), imgs);
    const uint32_t blocks = grid.x * grid.y * grid.z;
    auto reduce = [&](int l) {
        const uint32_t n = (uint32_t)(L_COUT[l] * 9 * L_CIN[l] + L_COUT[l]);
        hipLaunchKernelGGL(k_dn_reduce, dim3(cdiv(n, 256)), dim3(256), 0, d->stream, d->part, blocks, n,
                           d->g + L.w[l]);
    };
    // the last layer: dy (8 ch, 3 used) with input act[NL - 1]
    hipLaunchKernelGGL((k_dn_wgrad<HID, 3, IN_CH>), grid, dim3(256), 0, d->stream, d->act[NL - 1], d->dy8, W, H,
                       d->part);
    reduce(NL - 1);
    conv<IN_CH, 2, EPI_MASK>(d, d->dy8, W, H, imgs, d->bwd[NL - 1], nullptr, d->gA, d->act[NL - 1], nullptr, nullptr,
                             nullptr, nullptr, 0.0f, nullptr);
    uint16_t *cur = d->gA, *nxt = d->gB;
    for (int l = NL - 2; l >= 1; l--) {
        hipLaunchKernelGGL((k_dn_wgrad<HID, HID, HID>), grid, dim3(256), 0, d->stream, d->act[l], cur, W, H, d->part);
        reduce(l);
        conv<HID, 2, EPI_MASK>(d, cur, W, H, imgs, d->bwd[l], nullptr, nxt, d->act[l], nullptr, nullptr, nullptr,
                               nullptr, 0.0f, nullptr);
        uint16_t *t = cur;
        cur = nxt;
        nxt = t;
    }
    hipLaunchKernelGGL((k_dn_wgrad<IN_CH, HID, HID>), grid, dim3(256), 0, d->stream, d->act[0], cur, W, H, d->part);
    reduce(0);
}

}  // namespace

// ====================================================================== C ABI
uint64_t gps_denoiser_param_count(void) { return Layout().total; }

const char *gps_denoiser_last_error(const GpsDenoiser *d) { return d ? d->err.c_str() : ""; }

int gps_denoiser_create(int device, uint32_t seed, GpsDenoiser **out) {
    *out = nullptr;
    GpsDenoiser *d = new GpsDenoiser();
    d->device = device;
    auto bad = [&](const char *) {
        gps_denoiser_destroy(d);
        return -1;
    };
    if (hipSetDevice(device) != hipSuccess) return bad("device");
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) return bad("stream");
    const uint32_t n = d->lay.total;
    if (hipMalloc(&d->w, n * 4) != hipSuccess || hipMalloc(&d->m, n * 4) != hipSuccess ||
        hipMalloc(&d->v, n * 4) != hipSuccess || hipMalloc(&d->g, n * 4) != hipSuccess ||
        hipMalloc(&d->counts, 16) != hipSuccess)
        return bad("alloc");
    for (int l = 0; l < NL; l++) {
        // forward operands, then the biases (HID floats) the fused forward reads with them
        const size_t fw = (size_t)((L_COUT[l] + 15) / 16) * ksteps(L_CIN[l]) * 64 * 16 + HID * sizeof(float);
        const size_t bw = (size_t)((L_CIN[l] + 15) / 16) * ksteps(L_COUT[l] == 3 ? IN_CH : HID) * 64 * 16;
        if (hipMalloc(&d->fwd[l], fw) != hipSuccess || hipMalloc(&d->bwd[l], bw) != hipSuccess) return bad("alloc");
    }
    // He initialisation (normal, std sqrt(2 / fan_in) over the 7 used input channels for layer 0); the last
    // layer starts at a tenth of that so the initial output stays near the mean
    std::vector<float> h(n, 0.0f);
    Pcg rng(seed);
    for (int l = 0; l < NL; l++) {
        const int cin = L_CIN[l], cout = L_COUT[l];
        const int used = l == 0 ? 7 : cin;
        double sd = std::sqrt(2.0 / (9.0 * used));
        if (l == NL - 1) sd *= 0.1;
        for (int co = 0; co < cout; co++)
            for (int t = 0; t < 9; t++)
                for (int ci = 0; ci < cin; ci++)
                    h[d->lay.w[l] + ((size_t)co * 9 + t) * cin + ci] = ci < used ? (float)(sd * rng.normal()) : 0.0f;
    }
    if (gps_denoiser_set_params(d, h.data()) != 0) return bad("init");
    *out = d;
    return 0;
}

void gps_denoiser_destroy(GpsDenoiser *d) {
    if (!d) return;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    dnfree(d->w);
    dnfree(d->m);
    dnfree(d->v);
    dnfree(d->g);
    dnfree(d->counts);
    for (int l = 0; l < NL; l++) {
        dnfree(d->fwd[l]);
        dnfree(d->bwd[l]);
        dnfree(d->act[l]);
    }
    dnfree(d->dy8);
    dnfree(d->gA);
    dnfree(d->gB);
    dnfree(d->mean);
    dnfree(d->out);
    dnfree(d->tgt);
    dnfree(d->part);
    dnfree(d->lpart);
    dnfree(d->ratios);
    dnfree(d->px);
    dnfree(d->pmean);
    dnfree(d->ptgt);
    dnfree(d->bidx);
    dnfree(d->pweight);
    dnfree(d->bweight);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

int gps_denoiser_set_params(GpsDenoiser *d, const float *params) {
    DN_TRY(d, hipSetDevice(d->device));
    const uint32_t n = d->lay.total;
    DN_TRY(d, hipMemcpyAsync(d->w, params, n * 4, hipMemcpyHostToDevice, d->stream));
    DN_TRY(d, hipMemsetAsync(d->m, 0, n * 4, d->stream));
    DN_TRY(d, hipMemsetAsync(d->v, 0, n * 4, d->stream));
    d->adam_t = 0;
    if (pack_all(d) != 0) return -1;
    DN_TRY(d, hipStreamSynchronize(d->stream));
    return 0;
}

int gps_denoiser_get_params(GpsDenoiser *d, float *params) {
    DN_TRY(d, hipSetDevice(d->device));
    DN_TRY(d, hipMemcpyAsync(params, d->w, d->lay.total * 4, hipMemcpyDeviceToHost, d->stream));
    DN_TRY(d, hipStreamSynchronize(d->stream));
    return 0;
}

// features of renderer `r` into images [0] of the activation buffers
// `start`, if given, is recorded on the denoiser's stream once the renderer's work has finished
static int load_renderer(GpsDenoiser *d, GpsRenderer *r, uint32_t &W, uint32_t &H, hipEvent_t start = nullptr) {
    uint32_t spp_side = 1, deg = 0;
    uint64_t n = 0;
    gps_renderer_info(r, &W, &H, &spp_side, &n, &deg);
    const float4 *accum, *mom;
    const float *zmin;
    renderer_device_buffers(r, &accum, &mom, &zmin);
    if (!mom) return dfail(d, "the renderer's denoiser features are not enabled (gps_renderer_set_features)");
    const uint32_t X = W * H;
    if (ensure_act(d, X, 1, W, H) != 0) return -1;
    // the renderer's work is ordered before ours
    if (gps_renderer_wait_idle(r) != 0) return dfail(d, "renderer wait failed");
    if (start) (void)hipEventRecord(start, d->stream);
    hipLaunchKernelGGL(k_dn_features, dim3(cdiv(X, 256)), dim3(256), 0, d->stream, accum, mom, zmin, X,
                       (float)(spp_side * spp_side), d->act[0], d->mean, d->counts);
    return 0;
}

int gps_denoiser_apply(GpsDenoiser *d, GpsRenderer *r, float fallback_ratio, GpsDenoiseInfo *info) {
    DN_TRY(d, hipSetDevice(d->device));
    uint32_t W, H;
    hipEvent_t e0, e1;
    DN_TRY(d, hipEventCreate(&e0));
    DN_TRY(d, hipEventCreate(&e1));
    int rc = load_renderer(d, r, W, H, e0);
    if (rc == 0) {
        forward_infer(d, W, H);
        const dim3 tg(cdiv(W, FB_TILE), cdiv(H, FB_TILE));
        if (fallback_ratio > 0.0f)
            hipLaunchKernelGGL(k_dn_fallback, tg, dim3(256), 0, d->stream, d->out, d->mean, W, H, fallback_ratio,
                               d->counts, d->ratios);
        (void)hipEventRecord(e1, d->stream);
        uint32_t fb = 0;
        if (hipMemcpyAsync(&fb, d->counts, 4, hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
            hipStreamSynchronize(d->stream) != hipSuccess) {
            rc = dfail(d, "denoise failed");
        } else {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (info) {
                info->gpu_ms = ms;
                info->tiles = tg.x * tg.y;
                info->fallback_tiles = fallback_ratio > 0.0f ? fb : 0u;
                info->width = W;
                info->height = H;
            }
            d->last_w = W;
            d->last_h = H;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int gps_denoiser_set_fused(GpsDenoiser *d, int enable) {
    d->fused = enable != 0;
    return FUSED_FITS ? 0 : 1;
}

int gps_denoiser_read(GpsDenoiser *d, float *rgba) {
    if (!d->last_w) return dfail(d, "nothing denoised yet");
    DN_TRY(d, hipMemcpyAsync(rgba, d->out, (size_t)d->last_w * d->last_h * sizeof(float4), hipMemcpyDeviceToHost,
                             d->stream));
    DN_TRY(d, hipStreamSynchronize(d->stream));
    return 0;
}

int gps_denoiser_read_features(GpsDenoiser *d, float *x8, float *mean4) {
    if (!d->last_w) return dfail(d, "nothing denoised yet");
    const size_t X = (size_t)d->last_w * d->last_h;
    std::vector<uint16_t> h(X * IN_CH);
    DN_TRY(d, hipMemcpyAsync(h.data(), d->act[0], X * IN_CH * 2, hipMemcpyDeviceToHost, d->stream));
    if (mean4) DN_TRY(d, hipMemcpyAsync(mean4, d->mean, X * sizeof(float4), hipMemcpyDeviceToHost, d->stream));
    DN_TRY(d, hipStreamSynchronize(d->stream));
    if (x8)
        for (size_t i = 0; i < h.size(); i++) x8[i] = bf2f(h[i]);
    return 0;
}

// ---- training pairs
int gps_denoiser_add_pair(GpsDenoiser *d, GpsRenderer *input, GpsRenderer *target) {
    DN_TRY(d, hipSetDevice(d->device));
    uint32_t tw = 0, th = 0, ts = 0, tdeg = 0;
    uint64_t tn = 0;
    gps_renderer_info(target, &tw, &th, &ts, &tn, &tdeg);
    uint32_t W, H;
    if (load_renderer(d, input, W, H) != 0) return -1;
    if (tw != W || th != H) return dfail(d, "input and target renderers differ in size");
    if (d->pairs == 0 && d->pair_cap == 0) {
        d->pair_w = W;
        d->pair_h = H;
    }
    if (W != d->pair_w || H != d->pair_h) return dfail(d, "every training pair must have the same size");
    const size_t X = (size_t)W * H;
    if (d->pairs == d->pair_cap) {
        const uint32_t cap = d->pair_cap ? 2 * d->pair_cap : 64;
        uint16_t *px = nullptr;
        float4 *pm = nullptr, *pt = nullptr;
        float *pw = nullptr;
        DN_TRY(d, hipMalloc(&px, cap * X * IN_CH * 2));
        DN_TRY(d, hipMalloc(&pm, cap * X * sizeof(float4)));
        DN_TRY(d, hipMalloc(&pt, cap * X * sizeof(float4)));
        DN_TRY(d, hipMalloc(&pw, cap * sizeof(float)));
        if (d->pairs) {
            DN_TRY(d, hipMemcpyAsync(px, d->px, d->pairs * X * IN_CH * 2, hipMemcpyDeviceToDevice, d->stream));
            DN_TRY(d, hipMemcpyAsync(pm, d->pmean, d->pairs * X * sizeof(float4), hipMemcpyDeviceToDevice, d->stream));
            DN_TRY(d, hipMemcpyAsync(pt, d->ptgt, d->pairs * X * sizeof(float4), hipMemcpyDeviceToDevice, d->stream));
            DN_TRY(d, hipMemcpyAsync(pw, d->pweight, d->pairs * sizeof(float), hipMemcpyDeviceToDevice, d->stream));
            DN_TRY(d, hipStreamSynchronize(d->stream));
        }
        dnfree(d->px);
        dnfree(d->pmean);
        dnfree(d->ptgt);
        dnfree(d->pweight);
        d->px = px;
        d->pmean = pm;
        d->ptgt = pt;
        d->pweight = pw;
        d->pair_cap = cap;
    }
    const size_t o = (size_t)d->pairs * X;
    DN_TRY(d, hipMemcpyAsync(d->px + o * IN_CH, d->act[0], X * IN_CH * 2, hipMemcpyDeviceToDevice, d->stream));
    DN_TRY(d, hipMemcpyAsync(d->pmean + o, d->mean, X * sizeof(float4), hipMemcpyDeviceToDevice, d->stream));
    const float4 *tacc, *tm;
    const float *tz;
    renderer_device_buffers(target, &tacc, &tm, &tz);
    if (gps_renderer_wait_idle(target) != 0) return dfail(d, "target wait failed");
    hipLaunchKernelGGL(k_dn_mean_of, dim3(cdiv(X, 256)), dim3(256), 0, d->stream, tacc, (uint32_t)X, d->ptgt + o);
    // pair weight sqrt(ensembles / 4): leans the loss toward the many-ensemble inputs, whose errors are
    // small but decide where an image reaches a high PSNR, without letting them drown the few-ensemble ones
    const float wgt = std::sqrt((float)gps_renderer_accumulated(input) / 4.0f);
    DN_TRY(d, hipMemcpyAsync(d->pweight + d->pairs, &wgt, sizeof(float), hipMemcpyHostToDevice, d->stream));
    DN_TRY(d, hipGetLastError());
    DN_TRY(d, hipStreamSynchronize(d->stream));
    d->pairs++;
    return (int)(d->pairs - 1);
}

uint32_t gps_denoiser_pairs(const GpsDenoiser *d) { return d->pairs; }

int gps_denoiser_set_pair_weights(GpsDenoiser *d, const float *w) {
    if (d->pairs == 0) return 0;
    DN_TRY(d, hipMemcpyAsync(d->pweight, w, d->pairs * sizeof(float), hipMemcpyHostToDevice, d->stream));
    DN_TRY(d, hipStreamSynchronize(d->stream));
    return 0;
}

int gps_denoiser_pair_weights(GpsDenoiser *d, float *out) {
    if (d->pairs == 0) return 0;
    DN_TRY(d, hipMemcpyAsync(out, d->pweight, d->pairs * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    DN_TRY(d, hipStreamSynchronize(d->stream));
    return 0;
}

int gps_denoiser_clear_pairs(GpsDenoiser *d) {
    d->pairs = 0;
    return 0;
}

static int train_steps(GpsDenoiser *d, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float *losses,
                       bool update, float *grad_out) {
    DN_TRY(d, hipSetDevice(d->device));
    if (d->pairs == 0) return dfail(d, "no training pairs");
    if (batch == 0 || steps == 0) return dfail(d, "steps and batch must be > 0");
    const uint32_t W = d->pair_w, H = d->pair_h, X = W * H;
    if (ensure_act(d, X * batch, batch, W, H) != 0) return -1;
    if (batch > d->bidx_cap) {
        dnfree(d->bidx);
        dnfree(d->bweight);
        DN_TRY(d, hipMalloc(&d->bidx, batch * 4));
        DN_TRY(d, hipMalloc(&d->bweight, batch * 4));
        d->bidx_cap = batch;
    }
    float *dloss = nullptr;
    DN_TRY(d, hipMalloc(&dloss, steps * 4));
    Pcg rng(seed, 77u);
    std::vector<uint32_t> idx(batch);
    const dim3 grid(cdiv(W, TW), cdiv(H, TH), batch);
    const uint32_t blocks = grid.x * grid.y * grid.z;
    const float inv = 1.0f / (3.0f * (float)X * (float)batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int rc = 0;
    for (uint32_t s = 0; s < steps && rc == 0; s++) {
        for (uint32_t b = 0; b < batch; b++) idx[b] = update ? rng.next() % d->pairs : b % d->pairs;
        // the index upload is ordered on the stream; the host vector stays alive until the sync below
        if (hipMemcpyAsync(d->bidx, idx.data(), batch * 4, hipMemcpyHostToDevice, d->stream) != hipSuccess ||
            hipStreamSynchronize(d->stream) != hipSuccess) {
            rc = dfail(d, "batch upload failed");
            break;
        }
        hipLaunchKernelGGL(k_dn_gather, dim3(cdiv(X, 256), batch), dim3(256), 0, d->stream, d->px, d->pmean, d->ptgt,
                           d->pweight, d->bidx, X, d->act[0], d->mean, d->tgt, d->bweight);
        forward(d, W, H, batch, d->tgt, inv);
        hipLaunchKernelGGL(k_dn_sum, dim3(1), dim3(64), 0, d->stream, d->lpart, blocks, inv, dloss + s);
        backward(d, W, H, batch);
        if (update) {
            d->adam_t++;
            const float c1 = 1.0f - std::pow(b1, (float)d->adam_t), c2 = 1.0f - std::pow(b2, (float)d->adam_t);
            hipLaunchKernelGGL(k_dn_adam, dim3(cdiv(d->lay.total, 256)), dim3(256), 0, d->stream, d->w, d->g, d->m,
                               d->v, d->lay.total, lr, b1, b2, eps, c1, c2);
            if (pack_all(d) != 0) rc = -1;
        }
        if (hipGetLastError() != hipSuccess) rc = dfail(d, "training launch failed");
    }
    if (rc == 0 && losses) {
        if (hipMemcpyAsync(losses, dloss, steps * 4, hipMemcpyDeviceToHost, d->stream) != hipSuccess) rc = -1;
    }
    if (rc == 0 && grad_out) {
        if (hipMemcpyAsync(grad_out, d->g, d->lay.total * 4, hipMemcpyDeviceToHost, d->stream) != hipSuccess) rc = -1;
    }
    if (hipStreamSynchronize(d->stream) != hipSuccess && rc == 0) rc = dfail(d, "training failed");
    (void)hipFree(dloss);
    return rc;
}

int gps_denoiser_train(GpsDenoiser *d, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float *losses) {
    return train_steps(d, steps, batch, lr, seed, losses, true, nullptr);
}

int gps_denoiser_loss_grad(GpsDenoiser *d, uint32_t batch, float *loss, float *grad) {
    return train_steps(d, 1, batch, 0.0f, 0u, loss, false, grad);
}

int gps_denoiser_read_output(GpsDenoiser *d, uint32_t images, float *rgba) {
    const size_t n = (size_t)images * d->pair_w * d->pair_h;
    if (n == 0 || n > d->cap_px) return dfail(d, "no such output");
    DN_TRY(d, hipMemcpyAsync(rgba, d->out, n * sizeof(float4), hipMemcpyDeviceToHost, d->stream));
    DN_TRY(d, hipStreamSynchronize(d->stream));
    return 0;
}
