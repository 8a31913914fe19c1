"""Exhaustive checks of the specialised device math (gps_devmath.h: div_normal, div_by, sqrt_normal, exp_mid and the
branch-free quadrant selects of sincos_fin) against the general correctly rounded operations, bit for bit,
over every argument the particle kernels give them; and of the denoiser's hardware bf16 rounding.  The kernels are in tests/gpu_math_check.hip."""
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
    for name in ("mathcheck_sqrt", "mathcheck_div_log", "mathcheck_div_eol", "mathcheck_sincos", "mathcheck_bf16",
                 "mathcheck_exp_mid", "mathcheck_div_by"):
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


def test_exp_mid_equals_exp(lib):
    # the alpha blend's exp(-r2 / 2), r2 <= k^2 <= 2 ln 255: exp_mid (ldexp scaling, no range checks) against exp()
    # for every float in [-80, 0]
    assert _bad(lib.mathcheck_exp_mid) == 0


@pytest.mark.parametrize("extremes", [0, 1])
def test_div_by_shared_reciprocal(lib, extremes):
    # the alpha blend's u = N / |b1|^2 and v = N / |b2|^2 with the reciprocal refined once per splat: equal to the
    # correctly rounded quotient over 2^32 pairs in the guarded ranges (wherever |N| >= 2^-100)
    assert _bad(lib.mathcheck_div_by, ctypes.c_int(extremes)) == 0


def test_bf16_pack_rounds_to_nearest_even(lib):
    # the denoiser's forward activations: v_cvt_pk_bf16_f32 equals the integer rounding for every float but NaN
    assert _bad(lib.mathcheck_bf16) == 0


NAMES = ["E_R", "E_T", "ETA_TE", "TAU_E", "ETA_TF", "TAU_F", "ETA_HE", "ETA_HE_X", "ETA_HF", "ETA_HF_X", "E_RCP",
         "U_MAX", "DELTA_N"]


def test_count_only_error_constants_hold(lib):
    """Every error constant of the count-only path (gps_fastcount.h) is at least the maximum error measured
    exhaustively over its argument set, and DELTA_N / U_MAX follow from them."""
    c = (ctypes.c_double * len(NAMES))()
    assert lib.mathcheck_constants(c, len(NAMES)) == len(NAMES)
    k = dict(zip(NAMES, c))
    out = (ctypes.c_double * 3)()
    meas = {}
    for which in range(7):
        assert lib.mathcheck_measure(which, out) == 0
        meas[which] = tuple(out)
        print(f"measure {which}: {tuple(out)}", flush=True)
    eps = 2.0 ** -24
    r_max = meas[0][1]
    assert meas[0][0] <= k["E_R"] and meas[0][2] <= r_max * (1 + 1e-6)
    assert meas[1][0] <= k["E_T"]
    assert meas[2][0] <= k["ETA_TE"] and meas[2][1] <= k["TAU_E"]
    assert meas[3][0] <= k["ETA_TF"] and meas[3][1] <= k["TAU_F"]
    assert meas[4][0] <= k["ETA_HE"] and meas[4][1] <= k["ETA_HE_X"]
    assert meas[5][0] <= k["ETA_HF"] and meas[5][1] <= k["ETA_HF_X"]
    assert meas[6][0] <= k["E_RCP"]
    assert (r_max + k["E_R"]) * (1 + k["E_T"]) * (1 + eps) <= k["U_MAX"]
    assert r_max * k["E_T"] + k["E_R"] * (1 + k["E_T"]) + 2 * eps * r_max * (1 + k["E_T"]) <= k["DELTA_N"]
