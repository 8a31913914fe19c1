// Allocation-free exclusive u32 scan: gps_scan_exclusive_u32's kernels on caller-owned device buffers, with the
// look-back scratch (tile ticket and per-tile status words) in a file-scope device array instead of hipMalloc.
// For callers that cannot allocate device memory: code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc.  The function calls no hipMalloc, hipFree or allocating library routine.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gps_hip.h"
#include "gps_internal.h"

namespace {

constexpr uint64_t MAX_TILES = (GPS_SCAN_NOALLOC_MAX_N + SCAN_TILE - 1) / SCAN_TILE;
constexpr uint64_t STATUS_OFFSET = 32;   // u64 words before the status words: the ticket, padded to 256 bytes

// [0]: the tile ticket (low 32 bits); [32, 32 + MAX_TILES): look-back status words.  1 MiB + 256 B per device,
// held until the process ends.  In host code `scratch` names a host-side stand-in: kernels get the device
// address from hipGetSymbolAddress, resolved once per device.
static __device__ __align__(256) uint64_t scratch[STATUS_OFFSET + MAX_TILES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint64_t *g_ptr[MAX_DEVICES];   // scratch's device address on each device (null until resolved)

// scratch's address on `device`, which must be current; null on failure
uint64_t *scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        uint64_t *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

}  // namespace

int gps_scan_exclusive_u32_noalloc(int device, const uint32_t *in, uint32_t *out, uint64_t n, void *stream) {
    if (n == 0) return 0;
    if (n > GPS_SCAN_NOALLOC_MAX_N) return GPS_SCAN_NOALLOC_TOO_LARGE;
    if (!in || !out || device < 0 || device >= MAX_DEVICES) return -1;
    if (hipSetDevice(device) != hipSuccess) return -1;
    uint64_t *ptr = scratch_on(device);
    if (!ptr) return -1;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t tiles = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    // the previous call's ticket and status words: cleared by a fill kernel (no memset node under capture)
    launch_zero_u32(s, (uint32_t *)ptr, 2u * (STATUS_OFFSET + tiles));
    launch_scan_u32(s, in, out, n, (uint32_t *)ptr, ptr + STATUS_OFFSET, tiles, 1u);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
