"""Viewer policies (gps_hip/viewer.py; crates/gps-web/src/lib.rs:437-699) against a recording stand-in for the
renderer, and the orbit camera against crates/gps-core/src/camera.rs:94-145's tests. CPU only."""
import math
import types

import numpy as np
import pytest

gh = pytest.importorskip("gps_hip")
from gps_hip import viewer as V  # noqa: E402


class FakeRenderer:
    """Records calls; each rendered ensemble adds `demand` points (and `dropped`) to the counters.  Counter
    readbacks are asynchronous like the renderer's (gps.rs:1005-1038): one in flight, delivered at the second poll
    after its request."""

    def __init__(self, n=1000, demand=1e6, dropped=0.0, probe=0.0, skip_all=False):
        self.n, self.demand, self.dropped, self.probe, self.skip_all = n, demand, dropped, probe, skip_all
        self.calls = []
        self.acc = self.sub = 0
        self.points = self.drop = self.ens = 0
        self.probe_value = 0
        self.issued, self.pending, self.ready, self.latest = 0, None, False, self._snapshot(0)
        self.latest.ensembles = 0

    def reset_accum(self):
        self.calls.append(("reset_accum",))
        self.acc = self.sub = 0

    def reset_stats(self):
        self.points = self.drop = self.ens = 0

    def submitted(self):
        return self.sub

    def accumulated(self):
        return self.acc

    def gaussian_count(self):
        return self.n

    def sh_max_degree(self):
        return 0

    def resize(self, w, h, s):
        self.calls.append(("resize", w, h, s))

    def set_gaussians(self, g):
        self.calls.append(("set_gaussians", len(g)))

    def _snapshot(self, issue):
        return types.SimpleNamespace(points=self.points, dropped_points=self.drop, ensembles=self.ens,
                                     probe_demand=self.probe_value, visible_splats=0, truncated_splats=0,
                                     orphan_subpixels=0, candidates=0, undercovered_splats=0,
                                     skipped_ensembles=self.ens if self.skip_all else 0, issue=issue)

    def read_stats(self):
        return self._snapshot(0)

    def _poll(self):
        if self.pending is not None:
            if self.ready:
                self.latest, self.pending = self.pending, None
            else:
                self.ready = True

    def request_stats(self):
        self._poll()
        if self.pending is None:
            self.issued += 1
            self.pending, self.ready = self._snapshot(self.issued), False

    def latest_stats(self):
        self._poll()
        return self.latest

    def stats_issued(self):
        return self.issued

    def render_ensembles(self, cam, p, seed, n):
        self.calls.append(("render", seed, n, float(p.density_scale), int(p.max_points_per_splat)))
        self.sub += n
        self.ens += n
        if not self.skip_all:
            self.acc += n
        self.points += int(self.demand * p.density_scale * n)
        self.drop += int(self.dropped * p.density_scale * n)

    def render_sorted(self, cam, p, target):
        self.calls.append(("sorted", target))
        self.acc = self.sub = target

    def probe_demand(self, cam, p):
        self.calls.append(("probe",))
        self.probe_value = int(self.probe)

    def composite(self, p, w, h):
        return np.zeros((h, w, 4), np.uint8)


def renders(r):
    return [c for c in r.calls if c[0] == "render"]


# ------------------------------------------------------------------ orbit camera (camera.rs tests)
def test_eye_is_at_distance_and_target_is_centred():
    for up in [(0, 1, 0), (0, -1, 0), (0, 0, 1)]:
        c = V.OrbitCamera(target=(1.0, 2.0, 3.0), distance=7.0, up=up)
        c.orbit(0.7, 0.3)
        assert abs(np.linalg.norm(c.eye() - c.target) - 7.0) < 1e-12
        cam = c.to_camera(64, 48)
        R = np.array(cam.view_rot).reshape(3, 3)
        t = R @ (c.target - np.array(cam.view_pos))
        assert abs(t[0]) < 1e-9 and abs(t[1]) < 1e-9 and abs(t[2] - 7.0) < 1e-9


def test_pan_moves_the_scene_with_the_pointer():
    for up in [(0, 1, 0), (0, -1, 0)]:
        c = V.OrbitCamera(target=(0.3, -0.2, 1.0), distance=4.0, yaw=0.7, pitch=0.4, up=up)

        def project(c, p):
            cam = c.to_camera(1280, 720)
            R = np.array(cam.view_rot).reshape(3, 3)
            q = R @ (p - np.array(cam.view_pos))
            return cam.focal_x * q[0] / q[2] + 640.0, cam.focal_y * q[1] / q[2] + 360.0

        p0 = c.target.copy()
        before = project(c, p0)
        c.pan(40.0, -25.0, 720.0)
        after = project(c, p0)
        assert abs(after[0] - before[0] - 40.0) < 1e-6 and abs(after[1] - before[1] + 25.0) < 1e-6


def test_pitch_is_clamped_and_world_up_stays_up():
    c = V.OrbitCamera()
    c.orbit(0.0, 10.0)
    assert c.pitch <= V.PITCH_LIMIT
    cam = V.OrbitCamera(pitch=0.2).to_camera(64, 64)
    assert (np.array(cam.view_rot).reshape(3, 3) @ np.array([0.0, 1.0, 0.0]))[1] < 0.0


def test_zoom_clamps_and_frame_sphere_fills_the_view():
    c = V.OrbitCamera()
    c.zoom(1e-12)
    assert c.distance == 1e-4
    c.frame_sphere((1.0, 2.0, 3.0), 2.0)
    assert np.allclose(c.target, [1, 2, 3]) and abs(c.distance - 2.0 / math.sin(0.45) * 1.25) < 1e-12


# ------------------------------------------------------------------ ensemble LOD
def test_manual_lod_renders_per_frame_until_the_target_with_contiguous_seeds():
    r = FakeRenderer()
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=3, target=10, max_points=16384)
    t = 0.0
    for _ in range(8):
        t += 16.7
        v.frame(t)
    rs = renders(r)
    assert [c[2] for c in rs] == [3, 3, 3, 1]
    assert [c[1] for c in rs] == [1, 4, 7, 10]        # seeds continue the history
    assert r.accumulated() == 10
    assert v.frame(t + 16.7) is False                 # converged: nothing new to present


def test_per_frame_is_clamped_to_the_submission_limit():
    v = V.Viewer(64, 48, renderer=FakeRenderer())
    v.configure(64, 48, 1, 1.0, per_frame=99, target=100, max_points=16384)
    assert v.per_frame == V.MAX_ENSEMBLES_PER_FRAME


def test_adaptive_lod_one_while_moving_with_the_moving_cap_then_grows_when_still():
    r = FakeRenderer()
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=1, target=100000, max_points=16384)
    v.set_lod(True, 33.3, 512)
    t = 16.7
    v.frame(t)
    v.orbit(10.0, 0.0)                                # input: moving for the next 250 ms
    for _ in range(5):
        t += 16.7
        v.frame(t)
    moving = renders(r)[-5:]
    assert all(c[2] == 1 and c[4] == 512 for c in moving)
    n_before = len(r.calls)
    t += 300.0                                        # still again: the cap returns, so the history restarts
    v.frame(t)
    assert ("reset_accum",) in r.calls[n_before:]
    for _ in range(3):   # Auto path: the still view's demand probe is read back asynchronously (a frame or two)
        if renders(r)[-1][4] == 16384:
            break
        t += 16.7
        v.frame(t)
    assert renders(r)[-1][4] == 16384
    for _ in range(200):                              # fast frames: one more ensemble per 30 quiet frames
        t += 16.7
        v.frame(t)
    assert v.lod_n == V.MAX_ENSEMBLES_PER_FRAME
    for _ in range(40):                               # slow frames: back off at once
        t += 80.0
        v.frame(t)
    assert v.lod_n == 1


# ------------------------------------------------------------------ point budget
def test_point_budget_lowers_the_density_to_fit():
    r = FakeRenderer(demand=4.0e6, dropped=1.0e6)     # 5M points per ensemble at factor 1
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=1, target=1000, max_points=16384)
    v.set_point_budget(1.0)
    t = 0.0
    for _ in range(12):
        t += 16.7
        v.frame(t)
    assert abs(v.auto_density - 0.2) < 1e-6
    assert abs(renders(r)[-1][3] - 0.2) < 1e-6        # the density reaching the renderer
    for _ in range(20):                               # within e^0.25 of the want: no more restarts
        t += 16.7
        v.frame(t)
    assert abs(v.auto_density - 0.2) < 1e-6


def test_point_budget_factor_is_clamped_at_a_tenth():
    r = FakeRenderer(demand=1e9)
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=1, target=1000, max_points=16384)
    v.set_point_budget(1.0)
    for k in range(12):
        v.frame(16.7 * (k + 1))
    assert abs(v.auto_density - 0.1) < 1e-6


# ------------------------------------------------------------------ render path
def until_decided(v, t, dt=16.7):
    """Frames until the Auto path has decided the current history window: the probe's demand arrives with a
    readback issued after the probe, a frame or two later (lib.rs:578-600)."""
    v.frame(t)
    for _ in range(3):
        if v.decided_epoch == v.history_epoch:
            break
        t += dt
        v.frame(t)
    assert v.decided_epoch == v.history_epoch
    return t


def test_auto_path_takes_the_sorted_path_above_the_threshold_with_hysteresis():
    r = FakeRenderer(probe=50e6)
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=1, target=8, max_points=16384)
    v.set_render_path(2, 30.0)
    v.frame(16.7)
    assert not renders(r) and ("probe",) in r.calls   # waiting for the probe's readback: nothing rendered yet
    t = until_decided(v, 33.4)
    assert v.sorted_active and ("sorted", 8) in r.calls
    r.probe = 20e6                                    # above half the threshold: stays sorted
    v.reset()
    t = until_decided(v, t + 16.7)
    assert v.sorted_active
    r.probe = 10e6                                    # below half: back to GPS particles
    v.reset()
    until_decided(v, t + 16.7)
    assert not v.sorted_active and renders(r)


def test_auto_path_keeps_particles_for_pbvr_and_large_scenes():
    for method, n in ((gh.Method.Extinction, 1000), (gh.Method.Gps, V.SORT_MAX_SPLATS + 1)):
        r = FakeRenderer(n=n, probe=1e9)
        v = V.Viewer(64, 48, renderer=r)
        v.set_method(method, 3, True, False, True, 512.0, 0.01)
        v.frame(16.7)
        assert not v.sorted_active and not any(c[0] in ("probe", "sorted") for c in r.calls)


def test_slow_sorted_motion_hands_over_to_gps_particles():
    r = FakeRenderer(probe=50e6)
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=1, target=100000, max_points=16384)
    t = until_decided(v, 16.7)
    assert v.sorted_active
    for _ in range(40):                               # moving, 100 ms frames
        v.orbit(1.0, 0.0)
        t += 100.0
        v.frame(t)
    assert v.moving_prefers_gps and not v.sorted_active


def test_forced_paths():
    r = FakeRenderer(probe=0.0)
    v = V.Viewer(64, 48, renderer=r)
    v.set_render_path(1, 30.0)
    v.frame(16.7)
    assert v.sorted_active
    v.set_render_path(0, 30.0)
    v.frame(33.4)
    assert not v.sorted_active


def test_stops_submitting_when_every_ensemble_is_skipped():
    r = FakeRenderer(skip_all=True)
    v = V.Viewer(64, 48, renderer=r)
    v.configure(64, 48, 1, 1.0, per_frame=4, target=100, max_points=16384)
    v.set_render_path(0, 30.0)
    for k in range(10):
        v.frame(16.7 * (k + 1))
    assert r.submitted() == 8                          # two frames of 4, then stalled


def test_settings_clamp_like_the_viewer():
    v = V.Viewer(64, 48, renderer=FakeRenderer())
    v.set_lod(True, 1.0, 1)
    assert v.lod_target_ms == 8.0 and v.moving_cap == 64
    v.set_method(2, 3, True, False, True, 0.0, 1e9)
    assert v.params.pbvr.radial_correction == 1 and v.params.pbvr.base_k == 1.0
    assert abs(v.params.pbvr.reference_pixel_length - 1e3) < 1e-3
    v.set_method(1, 3, True, False, True, 512.0, 0.01)  # the radial correction is Extinction C3 only
    assert v.params.pbvr.radial_correction == 0
    with pytest.raises(ValueError):
        v.set_camera(0, 0, 0, -1.0, 0, 0, 0.9)
    v.set_camera(0, 0, 0, 3.0, 0.1, 9.0, 0.9)
    assert v.cam.pitch == V.PITCH_LIMIT


def test_demo_scene_and_framing():
    r = FakeRenderer()
    v = V.Viewer(64, 48, renderer=r)
    assert v.load_demo() == "600 Gaussians (demo)"
    g = V.demo_gaussians()
    assert g.shape == (600,) and np.isfinite(g["pos_opacity"]).all()
    radius = np.linalg.norm(g["pos_opacity"][:, :3], axis=1)
    assert radius.min() > 0.55 and radius.max() < 1.15   # three shells at 0.6, 0.85 and 1.1
    assert v.cam.pitch == 0.25 and v.cam.yaw == 0.6 and v.cam.up[1] == 1.0
