"""The 64-bit atomicMin microbenchmark library (gps_atomic_bench.hip), which measures the atomic ceiling that the
particle kernel's atomic rate is read against: argument checks on the host (no GPU call), one small run on the
GPU."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gps-pbvr-web_amd", "libgps_atomicbench.so")


def load():
    lib = ctypes.CDLL(LIB)
    lib.gps_atomic_bench.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    lib.gps_atomic_bench.restype = ctypes.c_int
    return lib


def test_atomic_bench_rejects_bad_arguments():
    lib = load()
    ms = ctypes.c_float()
    bad = [(9, 0, 0, 1, 1, 1), (31, 0, 0, 1, 1, 1), (20, 21, 0, 1, 1, 1), (20, 0, 2, 1, 1, 1), (20, 0, -1, 1, 1, 1),
           (20, 0, 0, 0, 1, 1), (20, 0, 0, 65537, 1, 1), (20, 0, 0, 1, 0, 1), (20, 0, 0, 1, 4097, 1),
           (20, 0, 0, 1, 1, 0), (20, 0, 0, 1, 1, 101)]
    for args in bad:
        assert lib.gps_atomic_bench(*args, ctypes.byref(ms)) == -1, args
    assert lib.gps_atomic_bench(20, 0, 0, 1, 1, 1, None) == -1


@pytest.mark.gpu
def test_atomic_bench_runs():
    lib = load()
    for mode in (0, 1):
        for log2_w in (0, 6, 16):
            ms = ctypes.c_float(-1.0)
            assert lib.gps_atomic_bench(16, log2_w, mode, 64, 16, 2, ctypes.byref(ms)) == 0
            assert ms.value > 0.0
