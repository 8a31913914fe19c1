"""Viewer policies over the renderer API: the interactive viewer's frame loop (crates/gps-web/src/lib.rs:86-716)
without its canvas.

- Ensemble LOD (`set_lod`, lib.rs:437-441, 643-666): Manual renders `per_frame` ensembles per frame; Adaptive
  renders one while the camera moves (with the per-splat cap lowered to `moving_cap`, a coarser preview whose
  history restarts once the camera is still) and, while still, backs off at once when the frame time exceeds
  the target and adds one ensemble per 30 quiet frames below 60% of it.
- Point budget (`set_point_budget`, lib.rs:447-453, 538-552): the density factor follows budget / demand, with
  demand = (points + dropped points) per ensemble at factor 1 from the current history's counters, changed
  only when it moves by more than e^0.25.
- Render path (`set_render_path`, lib.rs:461-471, 556-622): GPS particles, depth-sorted alpha, or Auto. Auto
  probes the GPS demand of a still view (`probe_demand`) and takes the sorted path above the threshold (0.5x the
  threshold to leave it again); a moving view keeps the last choice, and 20 moving sorted frames slower than
  60 ms (frame-time EMA) hand motion over to GPS particles. Only the GPS particle model uses the sorted path.

`frame(now_ms)` is one animation tick. Where the reference presents to its canvas, this composites into
`Viewer.image` (RGBA8, the output size). Counters are read back as the reference reads them: asynchronously, one
readback in flight (`request_stats` / `latest_stats` / `stats_issued`), and a probe's demand is taken only from a
readback issued after the probe.
"""
import json
import math

import numpy as np

from . import GAUSSIAN_DTYPE, GpsRenderer, Method, RenderParams, look_at, orbit_eye, robust_sphere

PITCH_LIMIT = 1.553                  # crates/gps-core/src/camera.rs:28
SORT_MAX_SPLATS = 8_000_000          # lib.rs:84
SORT_SLOW_MS = 60.0                  # lib.rs:86
MAX_ENSEMBLES_PER_FRAME = 4          # gps.rs:19 (MAX_ENSEMBLES_PER_SUBMIT), the LOD's upper end
U32 = 0xFFFFFFFF


class OrbitCamera:
    """crates/gps-core/src/camera.rs:8-91."""

    def __init__(self, target=(0.0, 0.0, 0.0), distance=5.0, yaw=0.0, pitch=0.0, up=(0.0, 1.0, 0.0), fov_y=0.9):
        self.target = np.asarray(target, np.float64).copy()
        self.distance, self.yaw, self.pitch, self.fov_y = float(distance), float(yaw), float(pitch), float(fov_y)
        self.up = np.asarray(up, np.float64).copy()

    def copy(self):
        return OrbitCamera(self.target, self.distance, self.yaw, self.pitch, self.up, self.fov_y)

    def eye(self):
        return orbit_eye(self.target, self.distance, self.yaw, self.pitch, self.up)

    def orbit(self, d_yaw, d_pitch):
        self.yaw += d_yaw
        self.pitch = min(max(self.pitch + d_pitch, -PITCH_LIMIT), PITCH_LIMIT)

    def pan(self, dx, dy, view_height):
        """camera.rs:50-56: the point under the target follows a drag of (dx, dy) pixels."""
        cam = self.to_camera(1, 1)
        right = np.array(cam.view_rot[0:3])
        down = np.array(cam.view_rot[3:6])
        s = self.distance * 2.0 * math.tan(0.5 * self.fov_y) / max(view_height, 1.0)
        self.target = self.target - (right * (dx * s) + down * (dy * s))

    def zoom(self, factor):
        self.distance = min(max(self.distance * factor, 1e-4), 1e6)

    def to_camera(self, w, h):
        return look_at(self.eye(), self.target, self.up, self.fov_y, w, h)

    def frame_sphere(self, centre, radius):
        self.target = np.asarray(centre, np.float64).copy()
        self.distance = (max(radius, 1e-6) / math.sin(0.5 * self.fov_y)) * 1.25


def _pcg(v):
    s = (v * 747796405 + 2891336453) & U32
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & U32
    return ((w >> 22) ^ w) & U32


def demo_gaussians():
    """The viewer's 600-splat demo scene (lib.rs:282-307): three spherical shells coloured by position, from
    the same pcg_hash stream (its cos / sin / ln evaluated in f64 and rounded to f32)."""
    g = np.zeros(600, GAUSSIAN_DTYPE)
    s = 12345

    def rnd():
        nonlocal s
        s = _pcg(s)
        return np.float32(s) * np.float32(1.0 / 4294967296.0)

    for i in range(600):
        u, v = rnd() * np.float32(2.0 * math.pi), rnd() * np.float32(2.0) - np.float32(1.0)
        ring = np.float32(0.6) + np.float32(0.25) * np.float32(i % 3)
        c = np.sqrt(np.float32(1.0) - v * v)
        x, y, z = ring * c * np.float32(math.cos(u)), v * ring, ring * c * np.float32(math.sin(u))
        sc = np.float32(math.log(np.float32(0.05) + np.float32(0.05) * rnd()))
        g[i]["pos_opacity"] = (x, y, z, np.float32(0.5) + np.float32(0.45) * rnd())
        g[i]["log_scale"] = (sc, sc + np.float32(0.3) * rnd(), sc - np.float32(0.3) * rnd(), 0.0)
        g[i]["rot_wxyz"] = (rnd() + np.float32(0.1), rnd() - np.float32(0.5), rnd() - np.float32(0.5),
                            rnd() - np.float32(0.5))
        g[i]["color"] = (np.float32(0.5) + np.float32(0.5) * x / np.float32(1.1),
                         np.float32(0.5) + np.float32(0.5) * y / np.float32(1.1),
                         np.float32(0.5) + np.float32(0.5) * z / np.float32(1.1), 0.0)
    return g


class Viewer:
    """The viewer state machine of lib.rs:86-148 (defaults as in `Viewer::create`, lib.rs:187-234) over a
    GpsRenderer (or any object with its methods)."""

    def __init__(self, width=1280, height=720, renderer=None):
        self.params = RenderParams()
        self.width, self.height = width, height
        self.output_width, self.output_height = width, height
        self.renderer = renderer if renderer is not None else GpsRenderer(width, height, self.params.spp_side)
        self.cam = OrbitCamera()
        self.home_cam = OrbitCamera()
        self.seed = self.next_seed = 1
        self.per_frame, self.target = 1, 256
        self.lod_adaptive, self.lod_target_ms = False, 33.3
        self.last_input_ms, self.last_frame_ms, self.ema_ms = -math.inf, 0.0, 16.7
        self.effective_per_frame, self.lod_n, self.lod_quiet = 1, 1, 0
        self.point_budget, self.auto_density = 0.0, 1.0
        self.path_mode, self.demand_threshold = 2, 30.0e6
        self.sorted_active, self.sorted_ok = False, True
        self.probe_epoch = self.decided_epoch = U32
        self.probe_wait, self.probe_issue, self.slow_frames = 0, 0, 0
        self.moving_prefers_gps, self.last_demand = False, 0
        self.frames_since_reset, self.ensemble_counter, self.history_epoch = 0, 0, 0
        self.moving_cap, self.lod_moving = 2048, False
        self.dirty, self.needs_present = True, True
        self.scene_info = ""
        self.image = None

    def invalidate(self):
        self.dirty = True

    # ---- scene
    def load_gaussians(self, gpu, sh_rest=None, sh_degree=0, y_down=True, info=None):
        """Load activated splats (`GpsRenderer.set_gaussians` layout) and frame them (lib.rs:260-280, 309-315)."""
        if len(gpu) == 0:
            raise ValueError("the file contains no Gaussians")
        if sh_degree:
            self.renderer.set_gaussians_sh(gpu, sh_rest, sh_degree)
        else:
            self.renderer.set_gaussians(gpu)
        self._frame_scene(gpu, y_down)
        self.params.srgb_output = 0
        self.needs_present = True
        note = f" (SH degree {sh_degree})" if sh_degree else ""
        self.scene_info = info or f"{len(gpu)} Gaussians{note}"
        self.invalidate()
        return self.scene_info

    def load_ply(self, source):
        from . import load_ply
        g = load_ply(source)
        return self.load_gaussians(g.to_gpu(), g.sh_rest, g.sh_degree)

    def load_demo(self):
        g = demo_gaussians()
        self.renderer.set_gaussians(g)
        self._frame_scene(g, False)
        self.scene_info = f"{len(g)} Gaussians (demo)"
        self.invalidate()
        return self.scene_info

    def _frame_scene(self, g, y_down):
        a = np.asarray(g)
        pos = (a["pos_opacity"][:, :3] if a.dtype == GAUSSIAN_DTYPE else
               np.asarray(a, np.float32).reshape(-1, 16)[:, :3]).astype(np.float64)
        sphere = robust_sphere(pos)
        c, radius = sphere if sphere is not None else (np.zeros(3), 1.0)
        self.cam = OrbitCamera(up=(0.0, -1.0, 0.0) if y_down else (0.0, 1.0, 0.0), pitch=0.25, yaw=0.6)
        self.cam.frame_sphere(c, max(radius, 1e-3))
        self.home_cam = self.cam.copy()

    def set_y_down(self, y_down):
        self.cam.up = np.array([0.0, -1.0, 0.0]) if y_down else np.array([0.0, 1.0, 0.0])
        self.home_cam.up = self.cam.up.copy()
        self.invalidate()

    # ---- settings (lib.rs:325-376, 396-471)
    def configure(self, width, height, spp_side, density, per_frame, target, max_points, divisor=1):
        out = (max(width, 1), max(height, 1))
        divisor = min(max(divisor, 1), 2)
        w, h = -(-out[0] // divisor), -(-out[1] // divisor)
        spp_side = min(max(spp_side, 1), 4)
        if (w, h) != (self.width, self.height) or spp_side != self.params.spp_side:
            self.renderer.resize(w, h, spp_side)
            self.width, self.height = w, h
        self.output_width, self.output_height = out
        self.params.spp_side = spp_side
        self.params.density_scale = min(max(density, 0.01), 100.0)
        self.params.max_points_per_splat = max(max_points, 1)
        self.per_frame = min(max(per_frame, 1), MAX_ENSEMBLES_PER_FRAME)
        self.target = max(target, 1)
        self.invalidate()

    def set_method(self, method, calibration, radial, centre_depth, jitter, base_k, reference_pixel_length):
        m = min(max(int(method), 0), 4)
        cal = min(max(int(calibration), 0), 3)
        self.params.method = m
        self.params.pbvr.calibration = cal
        self.params.pbvr.radial_correction = int(bool(radial) and m == Method.Extinction and cal == 3)
        self.params.pbvr.centre_depth = int(bool(centre_depth))
        self.params.pbvr.low_pass_jitter = int(bool(jitter))
        self.params.pbvr.base_k = min(max(base_k, 1.0), 1.0e6)
        self.params.pbvr.reference_pixel_length = min(max(reference_pixel_length, 1.0e-6), 1.0e3)
        self.invalidate()

    def set_seed(self, seed):
        self.seed = seed & U32
        self.invalidate()

    def set_sh_degree(self, degree):
        self.params.sh_degree = min(degree, 3)
        self.invalidate()

    def camera_json(self):
        c = self.cam
        return json.dumps({"target": [float(x) for x in c.target], "distance": c.distance, "yaw": c.yaw,
                           "pitch": c.pitch, "fov_y": c.fov_y, "up_y": -1 if c.up[1] < 0.0 else 1})

    def set_camera(self, tx, ty, tz, distance, yaw, pitch, fov_y):
        vals = (tx, ty, tz, distance, yaw, pitch, fov_y)
        if not all(math.isfinite(v) for v in vals) or distance <= 0.0 or not (0.05 <= fov_y < 3.0):
            raise ValueError("invalid camera values")
        self.cam.target = np.array([tx, ty, tz], np.float64)
        self.cam.distance = min(max(distance, 1e-4), 1e6)
        self.cam.yaw = yaw
        self.cam.pitch = min(max(pitch, -PITCH_LIMIT), PITCH_LIMIT)
        self.cam.fov_y = fov_y
        self.invalidate()

    def set_lod(self, adaptive, target_ms, moving_cap):
        self.lod_adaptive = bool(adaptive)
        self.lod_target_ms = min(max(target_ms, 8.0), 200.0)
        self.moving_cap = min(max(moving_cap, 64), 1 << 24)
        self.invalidate()

    def set_point_budget(self, millions):
        b = max(millions, 0.0) * 1.0e6 if math.isfinite(millions) else 0.0
        if b != self.point_budget:
            self.point_budget = b
            self.auto_density = 1.0
            self.invalidate()

    def set_render_path(self, mode, demand_threshold_millions):
        mode = min(int(mode), 2)
        thr = (min(max(demand_threshold_millions, 0.1), 100000.0) * 1.0e6
               if math.isfinite(demand_threshold_millions) else 30.0e6)
        if mode != self.path_mode or thr != self.demand_threshold:
            self.path_mode, self.demand_threshold = mode, thr
            self.sorted_ok, self.moving_prefers_gps, self.slow_frames = True, False, 0
            self.decided_epoch = U32
            self.invalidate()

    def set_srgb_output(self, on):
        self.params.srgb_output = int(bool(on))
        self.needs_present = True

    # ---- input (lib.rs:486-514)
    def orbit(self, dx, dy):
        self.last_input_ms = self.last_frame_ms
        self.cam.orbit(-dx * 0.005, dy * 0.005)
        self.invalidate()

    def pan(self, dx, dy, view_height):
        self.last_input_ms = self.last_frame_ms
        self.cam.pan(dx, dy, view_height)
        self.invalidate()

    def zoom(self, wheel):
        self.last_input_ms = self.last_frame_ms
        self.cam.zoom(math.exp(wheel * 0.001))
        self.invalidate()

    def reset(self):
        self.invalidate()

    def reset_view(self):
        self.cam = self.home_cam.copy()
        self.invalidate()

    def redraw(self):
        self.needs_present = True

    def accumulated(self):
        return self.renderer.accumulated()

    def _params(self, capped):
        p = self.params.copy(density_scale=self.params.density_scale * self.auto_density)
        if capped:
            p.max_points_per_splat = min(p.max_points_per_splat, self.moving_cap)
        return p

    # ---- one animation tick (lib.rs:519-699)
    def frame(self, now_ms):
        r = self.renderer
        dt = min(max(now_ms - self.last_frame_ms, 1.0), 500.0) if self.last_frame_ms > 0.0 else 16.7
        self.last_frame_ms = now_ms
        self.ema_ms += 0.1 * (dt - self.ema_ms)
        moving = self.lod_adaptive and now_ms - self.last_input_ms < 250.0
        if moving != self.lod_moving:
            self.lod_moving = moving
            self.dirty = True   # the particle cap changes, so the history restarts
        if self.dirty:
            r.reset_accum()
            r.reset_stats()
            self.history_epoch = (self.history_epoch + 1) & U32
            self.frames_since_reset = 0
            self.next_seed = self.seed
            self.dirty = False
            self.needs_present = True
        if r.submitted() <= self.target + 8:
            r.request_stats()
        acc = r.accumulated()
        # point budget: demand at density factor 1 from the counters of the current history window
        self.frames_since_reset = min(self.frames_since_reset + 1, U32)
        if (self.point_budget > 0.0 and not self.lod_moving and self.frames_since_reset >= 8
                and r.latest_stats().ensembles >= 2):
            st = r.latest_stats()
            demand = (st.points + st.dropped_points) / st.ensembles / self.auto_density
            if demand > 0.0:
                want = float(np.float32(min(max(self.point_budget / demand, 0.1), 1.0)))
                if abs(math.log(want / self.auto_density)) > 0.25:
                    self.auto_density = want
                    self.invalidate()
        # render path
        moving_now = now_ms - self.last_input_ms < 250.0
        cam = self.cam.to_camera(self.width, self.height)
        n_splats = r.gaussian_count()
        sorted_eligible = self.params.method == Method.Gps and 0 < n_splats <= SORT_MAX_SPLATS
        wait_for_probe = False
        if self.path_mode == 0:
            self.sorted_active = False
        elif self.path_mode == 1:
            self.sorted_active = sorted_eligible and self.sorted_ok
        else:
            if not (sorted_eligible and self.sorted_ok) or (moving_now and self.moving_prefers_gps):
                self.sorted_active = False
            elif not moving_now and self.decided_epoch != self.history_epoch:
                if self.probe_epoch != self.history_epoch:
                    try:
                        r.probe_demand(cam, self._params(False))
                        self.probe_epoch = self.history_epoch
                        self.probe_wait = 0
                        self.probe_issue = r.stats_issued()
                    except Exception:
                        self.sorted_ok = False
                        self.sorted_active = False
                if self.probe_epoch == self.history_epoch:
                    r.request_stats()
                    self.probe_wait += 1
                    # a readback issued after the probe (one in flight before it still shows the old window)
                    if r.latest_stats().issue > self.probe_issue:
                        demand = float(r.latest_stats().probe_demand)
                        self.last_demand = int(demand)
                        keep = 0.5 if self.sorted_active else 1.0   # hysteresis near the threshold
                        self.sorted_active = demand > keep * self.demand_threshold
                        self.decided_epoch = self.history_epoch
                    else:
                        wait_for_probe = True
        if self.sorted_active:
            if moving_now and self.path_mode == 2:
                self.slow_frames = self.slow_frames + 1 if self.ema_ms > SORT_SLOW_MS else 0
                if self.slow_frames >= 20:
                    self.moving_prefers_gps = True
                    self.sorted_active = False
                    self.slow_frames = 0
                    self.invalidate()
        else:
            self.slow_frames = 0

        stalled = acc == 0 and r.submitted() >= 8
        if self.sorted_active and not wait_for_probe:
            if acc < self.target:
                try:
                    r.render_sorted(cam, self._params(False), self.target)
                    self.effective_per_frame = 0
                    self.ensemble_counter = (self.ensemble_counter + 1) & U32
                    self.needs_present = True
                except Exception:
                    self.sorted_ok = False   # cannot hold the buffers: GPS particles take over
                    self.sorted_active = False
                    self.invalidate()
        elif wait_for_probe:
            pass
        elif acc < self.target and not stalled:
            if moving_now:
                self.lod_n, self.lod_quiet = 1, 0
            if not self.lod_adaptive:
                want = self.per_frame
            elif moving_now:
                want = 1
            else:
                self.lod_quiet += 1
                if self.ema_ms > self.lod_target_ms:
                    self.lod_n = max(self.lod_n - 1, 1)
                    self.lod_quiet = 0
                elif self.ema_ms < 0.6 * self.lod_target_ms and self.lod_quiet >= 30:
                    self.lod_n = min(self.lod_n + 1, MAX_ENSEMBLES_PER_FRAME)
                    self.lod_quiet = 0
                want = self.lod_n
            self.effective_per_frame = want
            n = min(want, self.target - acc)
            r.render_ensembles(cam, self._params(self.lod_moving), self.next_seed, n)
            self.next_seed = (self.next_seed + n) & U32
            self.ensemble_counter = (self.ensemble_counter + n) & U32
            self.needs_present = True
        if not self.needs_present:
            return False
        self.image = r.composite(self.params, self.output_width, self.output_height)
        self.needs_present = False
        return True

    def stats_json(self):
        """lib.rs:724-760 without the memory estimate."""
        r = self.renderer
        r.request_stats()
        s = r.latest_stats()
        names = {0: "GPS", 1: "PBVR Proportional", 2: "PBVR Extinction", 3: "PBVR ViewConditioned",
                 4: "Screen occupancy"}
        return json.dumps({
            "scene": self.scene_info.replace('"', "'"), "gaussians": r.gaussian_count(), "width": self.width,
            "height": self.height, "spp": self.params.spp_side ** 2, "accumulated": r.accumulated(),
            "target": self.target, "points": s.points, "visible": s.visible_splats,
            "truncated_splats": s.truncated_splats, "dropped_points": s.dropped_points, "seed": self.seed,
            "lod_moving": self.lod_moving, "ensembles_per_frame": self.effective_per_frame,
            "sh_max": r.sh_max_degree(), "sh_degree": min(self.params.sh_degree, r.sh_max_degree()),
            "orphan": s.orphan_subpixels, "candidates": s.candidates, "undercovered": s.undercovered_splats,
            "skipped": s.skipped_ensembles, "density_auto": round(self.auto_density, 3),
            "path": "sorted" if self.sorted_active else "gps", "demand": self.last_demand,
            "method": names.get(int(self.params.method), "?"), "yaw": round(self.cam.yaw, 4),
            "pitch": round(self.cam.pitch, 4), "distance": round(self.cam.distance, 4)})
