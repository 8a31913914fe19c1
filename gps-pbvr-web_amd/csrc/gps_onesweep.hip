// Onesweep LSD radix sort of 64-bit keys with 32-bit values, and the depth-sorted particle resolve.
//
// Sort (digits of 4 to 8 bits, the tuning's sort_digit_bits; stable):
//   k_os_hist   one read of the keys builds the histograms of every digit position the sort uses at once
//   k_os_pass   per digit pass, one kernel: ticketed tiles (16,384 keys by default) rank their keys with
//               wave64 ballots, publish per-digit counts and resolve their per-digit global offsets by decoupled
//               look-back
//               (one digit per thread), stage the tile in LDS in digit order, then write it out in
//               contiguous runs.  Passes whose digit is the same for every key are skipped.
//   Traffic per executed pass: 12 B read + 12 B written per key.  Narrower digits make cheaper passes (fewer
//   ballots per key, fewer look-back chains per tile) but more of them.
//
// Depth-sorted resolve (the cross-check of the packed-minimum path): each particle of an ensemble carries
// the key (ordered depth key << 32 | particle index) and its subpixel.  After the sort, the first sorted
// particle of each subpixel is its (depth key, particle index) minimum; particle indices grow with splat
// id (exclusive offsets), so that particle's (depth key, splat id) is exactly the packed minimum of the
// atomicMin path (gps_splat.wgsl:276-281, 526-531).
#include <string>
#include <vector>

#include "gps_internal.h"

namespace {

constexpr int OS_BLOCK = 256;
constexpr uint64_t OS_VAL_MASK = (1ull << 38) - 1ull;
constexpr uint32_t OS_EPOCH_MASK = (1u << 24) - 1u;
constexpr int OS_MAX_PASSES = 16;   // 64 bits of 4-bit digits

// histograms (256 words per pass) of the npass digits of rb bits from bit lo_bit up (the last may be narrower)
__global__ __launch_bounds__(OS_BLOCK) void k_os_hist(const uint64_t *__restrict__ keys, uint64_t n,
                                                      uint32_t *__restrict__ ghist, int lo_bit, int rb, int npass) {
    __shared__ uint32_t h[OS_MAX_PASSES][256];
    for (int i = threadIdx.x; i < npass * 256; i += OS_BLOCK) (&h[0][0])[i] = 0u;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * OS_BLOCK;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t mask = (1u << rb) - 1u;
    auto add = [&](uint64_t k, bool full) {   // full: every lane of this wave holds a key
        for (int d = 0; d < npass; d++) {
            uint32_t dg = (uint32_t)((k >> (lo_bit + rb * d)) & mask);
            // a wave whose 64 keys share the digit adds 64 once instead of 64 serialised LDS atomics
            uint32_t first = __builtin_amdgcn_readfirstlane(dg);
            if (full && __ballot(dg != first) == 0ull) {
                if (lane == 0) atomicAdd(&h[d][dg], 64u);
            } else {
                atomicAdd(&h[d][dg], 1u);
            }
        }
    };
    uint64_t i = (uint64_t)blockIdx.x * OS_BLOCK + threadIdx.x;
    // four loads in flight per thread while all four of the wave's keys are in range
    for (; i - lane + 3 * stride + 64 <= n; i += 4 * stride) {
        uint64_t kk[4];
#pragma unroll
        for (int u = 0; u < 4; u++) kk[u] = keys[i + (uint64_t)u * stride];
#pragma unroll
        for (int u = 0; u < 4; u++) add(kk[u], true);
    }
    for (; i < n; i += stride) add(keys[i], i - lane + 64 <= n);
    __syncthreads();
    for (int i = threadIdx.x; i < npass * 256; i += OS_BLOCK) {
        uint32_t c = (&h[0][0])[i];
        if (c) atomicAdd(&ghist[i], c);
    }
}

// exclusive scan of the values held one per thread over the block (s_tmp: one word per wave); every thread
// of the block calls it
__device__ __forceinline__ uint32_t scan_block_exclusive(uint32_t v, uint32_t *s_tmp) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) s_tmp[w] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (int i = 0; i < w; i++) base += s_tmp[i];
    __syncthreads();
    return base + inc - v;
}

// BLK threads x OS_ITEMS keys per thread per tile, RB-bit digits (the key's bits from `shift` up).  The tile is
// staged in digit order through one 8-byte-per-key LDS buffer, keys first and then values (a 4096-key tile takes
// 32 KB of LDS).  Threads 0..2^RB - 1 own one digit each for the per-digit counts, offsets and look-back.
template <int BLK, int OS_ITEMS, int RB>
__global__ __launch_bounds__(BLK) void k_os_pass(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                      uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                      uint64_t n, uint32_t shift, const uint32_t *__restrict__ ghist,
                                                      uint32_t *ticket, uint64_t *status, uint32_t epoch) {
    constexpr int WAVES = BLK / 64;
    constexpr int OS_TILE = BLK * OS_ITEMS;
    constexpr uint32_t NB = 1u << RB;   // digits; NB itself marks a key beyond n
    static_assert(RB >= 4 && RB <= 8 && (int)NB <= BLK, "digit width");
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tmp[WAVES];
    __shared__ uint32_t s_wh[WAVES][NB + 1];
    __shared__ uint32_t s_toff[NB];
    __shared__ uint64_t s_gpre[NB];
    __shared__ uint64_t s_keys[OS_TILE];
    uint32_t *s_vals = reinterpret_cast<uint32_t *>(s_keys);
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t d = lane; d < NB + 1u; d += 64) s_wh[w][d] = 0u;
    // global base of digit t for this pass
    const uint32_t gbase = scan_block_exclusive(t < NB ? ghist[t] : 0u, s_tmp);
    const uint64_t tile = s_tile;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t key[OS_ITEMS];
    uint32_t val[OS_ITEMS], rank[OS_ITEMS];  // rank, then the key's position in the staged tile (~0: none)
    // every load of the tile is issued before the ranking: the ranking's wave barriers would otherwise hold
    // each key's load back until the previous key is ranked
#pragma unroll
    for (int j = 0; j < OS_ITEMS; j++) {
        const uint64_t i = tile * OS_TILE + (uint64_t)w * (64 * OS_ITEMS) + (uint64_t)j * 64 + lane;
        key[j] = i < n ? kin[i] : 0ull;
        val[j] = i < n ? vin[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < OS_ITEMS; j++) {
        uint64_t i = tile * OS_TILE + (uint64_t)w * (64 * OS_ITEMS) + (uint64_t)j * 64 + lane;
        bool valid = i < n;
        uint32_t d = valid ? (uint32_t)((key[j] >> shift) & (NB - 1u)) : NB;
        uint64_t peers;
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        if (__ballot(d != d0) == 0ull) {
            peers = ~0ull;            // the whole wave shares one digit
        } else {
            peers = ~0ull;
#pragma unroll
            for (int b = 0; b <= RB; b++) {
                uint64_t bb = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? bb : ~bb;
            }
        }
        uint32_t before = s_wh[w][d];
        __builtin_amdgcn_wave_barrier();
        rank[j] = valid ? before + (uint32_t)__popcll(peers & lt) : 0xFFFFFFFFu;
        if ((peers & lt) == 0ull) s_wh[w][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit t: offsets of the waves inside the tile and the tile's count
    uint32_t run = 0;
    if (t < NB)
        for (int ww = 0; ww < WAVES; ww++) {
            uint32_t c = s_wh[ww][t];
            s_wh[ww][t] = run;
            run += c;
        }
    const uint32_t tcount = run;
    const uint32_t toff = scan_block_exclusive(tcount, s_tmp);
    if (t < NB) s_toff[t] = toff;
    // decoupled look-back for digit t (status words: 256 per tile, the first NB used)
    if (t < NB) {
        const uint64_t tag = (uint64_t)(epoch & OS_EPOCH_MASK) << 40;
        uint64_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&status[t], tag | (2ull << 38) | (uint64_t)tcount, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&status[tile * 256 + t], tag | (1ull << 38) | (uint64_t)tcount, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            int64_t tp = (int64_t)tile - 1;
            for (;;) {
                uint64_t wv;
                for (;;) {
                    wv = __hip_atomic_load(&status[(uint64_t)tp * 256 + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((uint32_t)(wv >> 40) == (epoch & OS_EPOCH_MASK) && ((wv >> 38) & 3ull) != 0ull) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                excl += wv & OS_VAL_MASK;
                if (((wv >> 38) & 3ull) == 2ull) break;
                tp--;
            }
            __hip_atomic_store(&status[tile * 256 + t], tag | (2ull << 38) | (excl + tcount), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_gpre[t] = (uint64_t)gbase + excl;
    }
    __syncthreads();
    // stage the tile's keys in digit order, write them out; then the values through the same buffer
    // (destinations fit 32 bits: the sort takes fewer than 2^32 keys)
#pragma unroll
    for (int j = 0; j < OS_ITEMS; j++) {
        if (rank[j] != 0xFFFFFFFFu) {
            const uint32_t d = (uint32_t)((key[j] >> shift) & (NB - 1u));
            rank[j] += s_toff[d] + s_wh[w][d];
            s_keys[rank[j]] = key[j];
        }
    }
    __syncthreads();
    const uint64_t tbase = tile * OS_TILE;
    const uint32_t cnt = (uint32_t)(n - tbase < (uint64_t)OS_TILE ? n - tbase : (uint64_t)OS_TILE);
    constexpr int QN = OS_ITEMS;
    uint32_t gdst[QN];
#pragma unroll
    for (int qi = 0; qi < QN; qi++) {
        const uint32_t q = t + (uint32_t)qi * BLK;
        if (q < cnt) {
            uint64_t k = s_keys[q];
            uint32_t d = (uint32_t)((k >> shift) & (NB - 1u));
            gdst[qi] = (uint32_t)(s_gpre[d] + (q - s_toff[d]));
            kout[gdst[qi]] = k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < OS_ITEMS; j++)
        if (rank[j] != 0xFFFFFFFFu) s_vals[rank[j]] = val[j];
    __syncthreads();
#pragma unroll
    for (int qi = 0; qi < QN; qi++) {
        const uint32_t q = t + (uint32_t)qi * BLK;
        if (q < cnt) vout[gdst[qi]] = s_vals[q];
    }
    if (t == 0 && tile * OS_TILE + OS_TILE >= n) *ticket = 0u;   // last tile: all tickets drawn
}

// first sorted position per subpixel
__global__ __launch_bounds__(256) void k_first_hit(const uint32_t *__restrict__ vals, uint64_t n,
                                                   uint32_t *__restrict__ first) {
    uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    uint32_t s = vals[q];
    if (s == 0xFFFFFFFFu) return;
    uint32_t seen = __hip_atomic_load(&first[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if ((uint32_t)q < seen) __hip_atomic_fetch_min(&first[s], (uint32_t)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// depth|id per subpixel from its first sorted particle: (depth key << 32) | owner splat of the particle index
__global__ __launch_bounds__(256) void k_build_depth(const uint64_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ first, uint64_t S,
                                                     const uint32_t *__restrict__ offsets, uint32_t N,
                                                     uint64_t *__restrict__ depth) {
    uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    uint32_t q = first[s];
    if (q == 0xFFFFFFFFu) {
        depth[s] = ~0ull;
        return;
    }
    uint64_t k = keys[q];
    uint32_t p = (uint32_t)k;
    uint32_t lo = 0, hi = N - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (offsets[mid] <= p) lo = mid; else hi = mid - 1u;
    }
    depth[s] = (k & 0xFFFFFFFF00000000ull) | (uint64_t)lo;
}

__global__ void k_u32_set(uint32_t *__restrict__ p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

unsigned nb(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ====================================================================== host
struct OnesweepState {
    uint32_t *ghist = nullptr;        // [OS_MAX_PASSES][256]
    uint32_t *ticket = nullptr;
    uint64_t *status = nullptr;
    uint64_t status_tiles = 0;
    uint32_t epoch = 1;
};

OnesweepState *onesweep_create() { return new OnesweepState(); }

void onesweep_destroy(OnesweepState *s) {
    if (!s) return;
    if (s->ghist) (void)hipFree(s->ghist);
    if (s->ticket) (void)hipFree(s->ticket);
    if (s->status) (void)hipFree(s->status);
    delete s;
}

using ProfList = std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>>;

namespace {
struct OProf {
    ProfList *l;
    hipStream_t st;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    OProf(ProfList *l_, hipStream_t s_, const char *n_) : l(l_), st(s_), name(n_) {
        if (l) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, st);
        }
    }
    ~OProf() {
        if (l) {
            (void)hipEventRecord(b, st);
            l->push_back({name, {a, b}});
        }
    }
};
}  // namespace

#define O_TRY(call)                                                                  \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return -1;                                                               \
        }                                                                            \
    } while (0)

// Stable sort of n (key, value) pairs by their bits lo_bit..63.  k[0]/v[0] hold the input; the result ends in
// buffer `result` (0 or 1).  Digits of the tuning's sort_digit_bits; reads the histograms back to skip passes
// whose digit is constant.  lo_bit = 32 skips the low word: valid when the input is already ordered by it (LSD
// passes over an already sorted field are identities).
template <int RB>
static void launch_pass(int items, unsigned tiles, hipStream_t st, const uint64_t *ki, const uint32_t *vi,
                        uint64_t *ko, uint32_t *vo, uint64_t n, uint32_t shift, const uint32_t *gh, uint32_t *ticket,
                        uint64_t *status, uint32_t epoch) {
    if (items == 8)
        hipLaunchKernelGGL((k_os_pass<256, 8, RB>), dim3(tiles), dim3(256), 0, st, ki, vi, ko, vo, n, shift, gh,
                           ticket, status, epoch);
    else if (items == 16)
        hipLaunchKernelGGL((k_os_pass<256, 16, RB>), dim3(tiles), dim3(256), 0, st, ki, vi, ko, vo, n, shift, gh,
                           ticket, status, epoch);
    else if (items == 32)
        hipLaunchKernelGGL((k_os_pass<512, 16, RB>), dim3(tiles), dim3(512), 0, st, ki, vi, ko, vo, n, shift, gh,
                           ticket, status, epoch);
    else
        hipLaunchKernelGGL((k_os_pass<1024, 16, RB>), dim3(tiles), dim3(1024), 0, st, ki, vi, ko, vo, n, shift, gh,
                           ticket, status, epoch);
}

int onesweep_sort(OnesweepState *s, hipStream_t st, uint64_t *k[2], uint32_t *v[2], uint64_t n, int &result,
                  std::string &err, ProfList *prof, int lo_bit) {
    result = 0;
    if (n <= 1) return 0;
    if (!s->ghist) {
        O_TRY(hipMalloc(&s->ghist, OS_MAX_PASSES * 256 * 4));
        O_TRY(hipMalloc(&s->ticket, 4));
        O_TRY(hipMemset(s->ticket, 0, 4));
    }
    // tiles of 256 x items keys; 32 runs as 512 threads x 16, 64 as 1024 threads x 16
    const GpsTuningDev tun = gps_tuning_current();
    const uint32_t ti = tun.sort_items_per_thread;
    const int items = (ti == 8 || ti == 16 || ti == 32) ? (int)ti : 64;
    const int rb = (tun.sort_digit_bits >= 4 && tun.sort_digit_bits <= 8) ? (int)tun.sort_digit_bits : 8;
    const int npass = (64 - lo_bit + rb - 1) / rb;
    const uint64_t tile_keys = (uint64_t)256 * items;
    uint64_t tiles = (n + tile_keys - 1) / tile_keys;
    if (tiles > s->status_tiles) {
        if (s->status) (void)hipFree(s->status);
        s->status = nullptr;
        O_TRY(hipMalloc(&s->status, tiles * 256 * 8));
        O_TRY(hipMemset(s->status, 0, tiles * 256 * 8));
        s->status_tiles = tiles;
    }
    hipLaunchKernelGGL(k_u32_set, dim3(8), dim3(256), 0, st, s->ghist, (uint64_t)npass * 256, 0u);
    {
        OProf p(prof, st, "sort_hist");
        unsigned grid = nb(n, OS_BLOCK * 16);
        if (grid > 4096) grid = 4096;
        hipLaunchKernelGGL(k_os_hist, dim3(grid), dim3(OS_BLOCK), 0, st, k[0], n, s->ghist, lo_bit, rb, npass);
    }
    std::vector<uint32_t> h((size_t)npass * 256);
    O_TRY(hipMemcpyAsync(h.data(), s->ghist, h.size() * 4, hipMemcpyDeviceToHost, st));
    O_TRY(hipStreamSynchronize(st));
    int cur = 0;
    for (int d = 0; d < npass; d++) {
        bool constant = false;
        for (int b = 0; b < 256; b++)
            if (h[(size_t)d * 256 + b] == n) constant = true;
        if (constant) continue;
        OProf p(prof, st, "sort_pass");
        const uint32_t shift = (uint32_t)(lo_bit + rb * d);
        const uint32_t *gh = s->ghist + 256 * d;
        if (rb == 4) launch_pass<4>(items, (unsigned)tiles, st, k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], n, shift, gh,
                                    s->ticket, s->status, s->epoch);
        else if (rb == 5) launch_pass<5>(items, (unsigned)tiles, st, k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], n, shift,
                                         gh, s->ticket, s->status, s->epoch);
        else if (rb == 6) launch_pass<6>(items, (unsigned)tiles, st, k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], n, shift,
                                         gh, s->ticket, s->status, s->epoch);
        else if (rb == 7) launch_pass<7>(items, (unsigned)tiles, st, k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], n, shift,
                                         gh, s->ticket, s->status, s->epoch);
        else launch_pass<8>(items, (unsigned)tiles, st, k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], n, shift, gh,
                            s->ticket, s->status, s->epoch);
        s->epoch = (s->epoch + 1u) & OS_EPOCH_MASK;
        if (!s->epoch) s->epoch = 1;
        cur ^= 1;
    }
    O_TRY(hipGetLastError());
    result = cur;
    return 0;
}

void launch_first_hit(hipStream_t s, const uint32_t *vals, uint64_t n, uint32_t *first, uint64_t S) {
    unsigned g = nb(S, 256);
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_u32_set, dim3(g), dim3(256), 0, s, first, S, 0xFFFFFFFFu);
    if (n) hipLaunchKernelGGL(k_first_hit, dim3(nb(n, 256)), dim3(256), 0, s, vals, n, first);
}

void launch_build_depth(hipStream_t s, const uint64_t *keys, const uint32_t *first, uint64_t S,
                        const uint32_t *offsets, uint32_t N, uint64_t *depth) {
    hipLaunchKernelGGL(k_build_depth, dim3(nb(S, 256)), dim3(256), 0, s, keys, first, S, offsets, N, depth);
}

// standalone stable sort (test hook): in place on host arrays
extern "C" int gps_sort_pairs_u64(int device, uint64_t *keys, uint32_t *vals, uint64_t n) {
    if (n == 0) return 0;
    if (n >= (1ull << 32)) return -1;
    if (hipSetDevice(device) != hipSuccess) return -1;
    std::string err;
    OnesweepState *s = onesweep_create();
    uint64_t *k[2] = {nullptr, nullptr};
    uint32_t *v[2] = {nullptr, nullptr};
    int rc = 0, res = 0;
    for (int i = 0; i < 2 && rc == 0; i++)
        if (hipMalloc(&k[i], n * 8) != hipSuccess || hipMalloc(&v[i], n * 4) != hipSuccess) rc = -1;
    if (rc == 0 && (hipMemcpy(k[0], keys, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(v[0], vals, n * 4, hipMemcpyHostToDevice) != hipSuccess))
        rc = -1;
    if (rc == 0) rc = onesweep_sort(s, nullptr, k, v, n, res, err, nullptr, 0);
    if (rc == 0 && (hipMemcpy(keys, k[res], n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(vals, v[res], n * 4, hipMemcpyDeviceToHost) != hipSuccess))
        rc = -1;
    for (int i = 0; i < 2; i++) {
        if (k[i]) (void)hipFree(k[i]);
        if (v[i]) (void)hipFree(v[i]);
    }
    onesweep_destroy(s);
    return rc;
}
