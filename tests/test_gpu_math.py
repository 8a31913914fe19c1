"""Exhaustive checks of the specialised device math (gps_devmath.h: div_normal, sqrt_normal and the
branch-free quadrant selects of sincos_fin) against the general correctly rounded operations, bit for bit,
over every argument the particle kernels give them.  The kernels are in tests/gpu_math_check.hip."""
import ctypes
import os
import struct

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgps_mathcheck.so")


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB)
    for name in ("mathcheck_sqrt", "mathcheck_div_log", "mathcheck_div_eol", "mathcheck_sincos"):
        getattr(lib, name).restype = ctypes.c_int
    return lib


def _bad(fn, *args):
    n = ctypes.c_ulonglong(123)
    assert fn(*args, ctypes.byref(n)) == 0
    return n.value


def test_sqrt_normal_box_muller_range(lib):
    # Box-Muller: -2 log u1 for u1 in [5.96e-8, 1 - 2^-24] lies in [1.19e-7, 33.3]; checked over [2^-24, 64]
    assert _bad(lib.mathcheck_sqrt, ctypes.c_uint32(fbits(2.0 ** -24)), ctypes.c_uint32(fbits(64.0))) == 0


def test_sqrt_normal_whole_stated_domain(lib):
    assert _bad(lib.mathcheck_sqrt, ctypes.c_uint32(fbits(2.0 ** -96)), ctypes.c_uint32(fbits(2.0 ** 126))) == 0


def test_div_normal_log_quotient(lib):
    assert _bad(lib.mathcheck_div_log) == 0


def test_div_normal_extinction_quotient(lib):
    assert _bad(lib.mathcheck_div_eol) == 0


def test_sincos_selects(lib):
    # the particle kernels take sin/cos of 2 pi u, u in [0, 1): every float in [0, 6.2832]
    assert _bad(lib.mathcheck_sincos, ctypes.c_uint32(fbits(6.2832))) == 0
