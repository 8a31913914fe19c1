"""CPU-only checks of the C ABI library: it loads, exports every function include/gps_hip.h declares,
and its host-side logic (defaults, memory estimate, radius table) matches the reference."""
import ctypes
import os
import re

import numpy as np

import gps_hip as gh
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gps_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(gps_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(gh.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(gh.EXPORTED) == names


def test_render_params_default_matches_reference():
    """RenderParams::default() / PbvrParams::default(), gps.rs:181-229"""
    p = gh.RenderParams()
    assert p.spp_side == 2 and p.density_scale == 1.0 and p.max_points_per_splat == 65536
    assert abs(p.near - 0.05) < 1e-7 and list(p.background) == [np.float32(0.05), np.float32(0.05), np.float32(0.08)]
    assert p.exposure == 1.0 and p.srgb_output == 1 and p.sh_degree == 3 and p.exact_radius == 0
    assert p.path == gh.RenderPath.Particle and p.method == gh.Method.Gps
    assert p.pbvr.calibration == 3 and p.pbvr.radial_correction == 0 and p.pbvr.centre_depth == 0
    assert p.pbvr.low_pass_jitter == 1 and p.pbvr.base_k == 512.0 and abs(p.pbvr.reference_pixel_length - 0.01) < 1e-9


def test_struct_sizes():
    """gps.rs:1099-1103 (GpuGaussian is 64 B) and the ABI struct layouts shared with the oracle."""
    assert gh.GAUSSIAN_DTYPE.itemsize == 64
    assert ctypes.sizeof(gh.RenderParams) == ctypes.sizeof(O.MParams)
    assert ctypes.sizeof(gh.GpsCamera) == ctypes.sizeof(O.OrcCamera) == 12 * 8 + 5 * 8


def test_memory_estimate_matches_plan():
    """gps.rs:1105-1110"""
    m = gh.estimate_memory(1920, 1080, 2, 0)
    assert m["depth_winner_bytes"] + m["accum_bytes"] == 8 * 1920 * 1080 * 4 + 16 * 1920 * 1080
    assert abs((m["depth_winner_bytes"] + m["accum_bytes"]) / 1048576.0 - 95.0) < 1.0


def test_radius_lut_equals_oracle_table():
    """The library's f64 radius table (lib.rs:624-636) is bit-identical to the oracle's restatement."""
    a = gh.radius_lut()
    b = O.radius_lut()
    assert a.shape == (65 * 129,) and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_cameras_match_oracle():
    c1 = gh.look_at((3.0, 2.0, -5.0), (0.0, 0.5, 1.0), (0.0, 1.0, 0.0), 0.9, 128, 96)
    c2 = O.look_at((3.0, 2.0, -5.0), (0.0, 0.5, 1.0), (0.0, 1.0, 0.0), 0.9, 128, 96)
    assert np.allclose(list(c1.view_rot), list(c2.view_rot), atol=1e-15)
    assert abs(c1.focal_y - c2.focal_y) < 1e-12
    o1 = gh.orbit_camera((1.0, 2.0, 3.0), 7.0, 0.7, 0.3, (0.0, -1.0, 0.0), 0.9, 64, 48)
    o2 = O.orbit_camera((1.0, 2.0, 3.0), 7.0, 0.7, 0.3, (0.0, -1.0, 0.0), 0.9, 64, 48)
    assert np.allclose(list(o1.view_pos), list(o2.view_pos), atol=1e-12)
    assert np.allclose(list(o1.view_rot), list(o2.view_rot), atol=1e-12)
