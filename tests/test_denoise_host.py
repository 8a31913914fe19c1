"""Host-side logic of the denoiser tooling (no GPU): window cameras, PSNR, parameter layout."""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gps-pbvr-web_amd"))
import gps_hip as gh  # noqa: E402
from gps_hip import denoise as D  # noqa: E402
import denoise_ref as R  # noqa: E402


def project(cam, p):
    rot = np.array(cam.view_rot).reshape(3, 3)
    q = rot @ (np.asarray(p) - np.array(cam.view_pos))
    return cam.focal_x * q[0] / q[2] + cam.cx, cam.focal_y * q[1] / q[2] + cam.cy


def test_window_camera_shifts_pixels():
    rng = np.random.default_rng(3)
    full = D.random_view(rng)
    win = D.window_camera(full, 300, 200)
    for p in ([0.1, 0.2, 0.3], [-1.0, 0.5, -2.0]):
        u, v = project(full, p)
        uw, vw = project(win, p)
        assert abs((u - 300) - uw) < 1e-9 and abs((v - 200) - vw) < 1e-9
    z = D.window_camera(full, full.cx, full.cy)
    assert not (z.cx == 0.0 and z.cy == 0.0)


def test_random_views_in_range():
    rng = np.random.default_rng(0)
    for _ in range(50):
        c = D.random_view(rng)
        d = math.sqrt(sum(x * x for x in c.view_pos))
        assert 2.0 - 1e-9 <= d <= 15.0 + 1e-9


def test_psnr_matches_oracle_definition():
    a = np.zeros((4, 4, 3))
    b = np.full((4, 4, 3), 0.1)
    assert abs(D.psnr(a, b) - 20.0) < 1e-9 and D.psnr(a, a) == 120.0


def test_parameter_layout():
    n = sum(co * 9 * ci + co for ci, co in R.LAYERS)
    assert gh.denoiser_param_count() == n == 30947
    v = np.arange(n, dtype=np.float32)
    parts = gh.denoiser_split(v)
    assert [w.shape for w, _ in parts] == [(co, 3, 3, ci) for ci, co in R.LAYERS]
    assert parts[0][1][0] == 32 * 9 * 8 and parts[-1][1][-1] == n - 1


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3], np.float32)
    y = R.bf16_round(x)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 + 2 ** -6 and y[3] == 1.0
    assert abs(y[4] - x[4]) <= abs(x[4]) * 2 ** -8
