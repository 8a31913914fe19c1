"""gps_scan_exclusive_u32_noalloc (the allocation-free scan) against gps_scan_exclusive_u32, through the C ABI.

The caller owns the device buffers and the stream (allocated here through the HIP runtime the library itself links,
libamdhip64.so.7); the entry point's own scratch is a file-scope device array shared by every call on the GPU.  Four
calls in a row, the later ones smaller than the first, so that a stale ticket or status word of an earlier call would
corrupt a later one: bit-exact against the allocating entry point and against the exclusive sum modulo 2^32."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gh = pytest.importorskip("gps_hip")

H2D, D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs, from the runtime libgps_hip.so links."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]), ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipStreamCreate", [C.POINTER(C.c_void_p)]), ("hipStreamDestroy", [C.c_void_p]),
                           ("hipStreamSynchronize", [C.c_void_p]), ("hipSetDevice", [C.c_int])):
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = C.c_int

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), nbytes))
        return p


def _has_gpu():
    try:
        return gh.lib.gps_renderer_create is not None and Hip().lib.hipSetDevice(0) == 0
    except OSError:
        return False


def _noalloc(hip, x, stream):
    nb = x.size * 4
    din, dout = hip.malloc(nb), hip.malloc(nb)
    try:
        hip.check(hip.lib.hipMemcpy(din, x.ctypes.data, nb, H2D))
        junk = np.full(x.size, 0xFFFFFFFF, np.uint32)
        hip.check(hip.lib.hipMemcpy(dout, junk.ctypes.data, nb, H2D))
        assert gh.lib.gps_scan_exclusive_u32_noalloc(0, din, dout, x.size, stream) == 0
        hip.check(hip.lib.hipStreamSynchronize(stream))
        out = np.empty_like(x)
        hip.check(hip.lib.hipMemcpy(out.ctypes.data, dout, nb, D2H))
        return out
    finally:
        hip.check(hip.lib.hipFree(din))
        hip.check(hip.lib.hipFree(dout))


def test_noalloc_scan_equals_allocating_scan():
    if not _has_gpu():
        pytest.skip("no GPU")
    hip = Hip()
    stream = C.c_void_p()
    hip.check(hip.lib.hipStreamCreate(C.byref(stream)))
    rng = np.random.default_rng(11)
    try:
        for n, hi in ((3_000_017, 1 << 32), (100_003, 7), (2049, 1 << 32), (1, 5)):
            print(f"[no-alloc scan] n={n} values < {hi}", flush=True)
            x = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
            ref = gh.scan_exclusive_u32(x)
            want = np.concatenate([np.zeros(1, np.uint64), np.cumsum(x[:-1], dtype=np.uint64)]) & np.uint64(0xFFFFFFFF)
            assert np.array_equal(ref, want.astype(np.uint32))
            assert np.array_equal(_noalloc(hip, x, stream), ref), n
    finally:
        hip.check(hip.lib.hipStreamDestroy(stream))


def test_noalloc_scan_rejects_sizes_above_its_scratch():
    if not _has_gpu():
        pytest.skip("no GPU")
    hip = Hip()
    buf = hip.malloc(64)
    try:
        zeros = np.zeros(16, np.uint32)
        hip.check(hip.lib.hipMemcpy(buf, zeros.ctypes.data, 64, H2D))
        too_large = int(gh.SCAN_NOALLOC_MAX_N) + 1
        assert gh.lib.gps_scan_exclusive_u32_noalloc(0, buf, buf, too_large, None) == gh.SCAN_NOALLOC_TOO_LARGE
        assert gh.lib.gps_scan_exclusive_u32_noalloc(0, buf, buf, 0, None) == 0
        out = np.ones(16, np.uint32)
        hip.check(hip.lib.hipMemcpy(out.ctypes.data, buf, 64, D2H))
        assert not out.any()   # nothing was written
    finally:
        hip.check(hip.lib.hipFree(buf))
