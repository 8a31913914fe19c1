"""Python binding of libgps_hip.so (the C ABI of include/gps_hip.h).

``GpsRenderer`` mirrors gps_render::gps::GpsRenderer (crates/gps-render/src/gps.rs:302-1065): the
same method names, argument meaning and error behaviour (a failed call raises ``GpsError`` carrying the
library's message, the reference's ``Err(String)``).  The renderer runs on the MI355X through the HIP
kernels only; there is no CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes as C
import enum
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libgps_hip.so")


class GpsError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise GpsError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       f"(make -C gps-pbvr-web_amd/csrc)")
    return C.CDLL(LIB_PATH)


lib = _load()

u8, u32, u64, f32, f64 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_float, C.c_double
P = C.POINTER

# GpuGaussian, gps.rs:26-38 (64 bytes)
GAUSSIAN_DTYPE = np.dtype([("pos_opacity", "<f4", 4), ("log_scale", "<f4", 4), ("rot_wxyz", "<f4", 4),
                           ("color", "<f4", 4)])
SH_C0 = 0.28209479177387814


class Method(enum.IntEnum):
    """gps.rs:104-140"""
    Gps = 0
    Proportional = 1
    Extinction = 2
    ViewConditioned = 3
    ScreenOccupancy = 4


class Calibration(enum.IntEnum):
    """gps.rs:144-154"""
    None_ = 0
    ObjectZoom = 1
    PerSplatDepth = 2
    PerSplatFootprint = 3


class RenderPath(enum.IntEnum):
    """gps.rs:93-100; ParticleSorted: depth-sorted resolve of the particle path (same bits as Particle)"""
    Primitive = 0
    Particle = 1
    ParticleSorted = 2


class GpsCamera(C.Structure):
    """OracleCamera, oracle.rs:33-42 (view_rot row-major)."""
    _fields_ = [("view_rot", f64 * 9), ("view_pos", f64 * 3), ("focal_x", f64), ("focal_y", f64), ("cx", f64),
                ("cy", f64), ("low_pass", f64)]


class PbvrParams(C.Structure):
    """gps.rs:169-179"""
    _fields_ = [("calibration", u32), ("radial_correction", u32), ("centre_depth", u32), ("low_pass_jitter", u32),
                ("base_k", f32), ("reference_pixel_length", f32)]


class RenderParams(C.Structure):
    """gps.rs:189-229; RenderParams() gives RenderParams::default()."""
    _fields_ = [("spp_side", u32), ("density_scale", f32), ("max_points_per_splat", u32), ("near", f32),
                ("background", f32 * 3), ("exposure", f32), ("srgb_output", u32), ("exact_radius", u32),
                ("sh_degree", u32), ("path", u32), ("method", u32), ("pbvr", PbvrParams)]

    def __init__(self, **kw):
        super().__init__()
        lib.gps_render_params_default(C.byref(self))
        for k, v in kw.items():
            if k == "pbvr" and isinstance(v, dict):
                for kk, vv in v.items():
                    setattr(self.pbvr, kk, int(vv) if isinstance(vv, (bool, enum.IntEnum)) else vv)
            elif k == "background":
                self.background[:] = list(v)
            else:
                setattr(self, k, int(v) if isinstance(v, (bool, enum.IntEnum)) else v)

    def copy(self, **kw):
        q = RenderParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(RenderParams))
        for k, v in kw.items():
            if k == "pbvr" and isinstance(v, dict):
                for kk, vv in v.items():
                    setattr(q.pbvr, kk, int(vv) if isinstance(vv, (bool, enum.IntEnum)) else vv)
            elif k == "background":
                q.background[:] = list(v)
            else:
                setattr(q, k, int(v) if isinstance(v, (bool, enum.IntEnum)) else v)
        return q


class GpuStats(C.Structure):
    """gps.rs:233-252"""
    _fields_ = [("points", u64), ("truncated_splats", u64), ("dropped_points", u64), ("visible_splats", u64),
                ("orphan_subpixels", u64), ("candidates", u64), ("undercovered_splats", u64),
                ("skipped_ensembles", u64), ("ensembles", u64), ("probe_demand", u64), ("issue", u32),
                ("reserved", u32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}

    def __eq__(self, other):
        return self.as_dict() == other.as_dict()


class Tuning(C.Structure):
    """Launch tuning (include/gps_hip.h GpsTuning); rendered bits never depend on it."""
    _fields_ = [("particle_waves_per_cu", u32), ("particles_per_thread", u32), ("sort_items_per_thread", u32),
                ("batch_ensembles", u32), ("depth_window_mib", u32), ("count_only", u32), ("particle_block", u32),
                ("sort_digit_bits", u32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchItem(C.Structure):
    _fields_ = [("renderer", C.c_void_p), ("camera", C.c_void_p), ("params", C.c_void_p), ("first_seed", u32),
                ("count", u32)]


class DenoiseInfo(C.Structure):
    _fields_ = [("gpu_ms", f64), ("width", u32), ("height", u32), ("tiles", u32), ("fallback_tiles", u32)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", u64), ("total_ms", f64)]


def _sig(name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


_R = C.c_void_p
_sig("gps_render_params_default", None, [P(RenderParams)])
_sig("gps_renderer_create", C.c_int, [u32, u32, u32, C.c_int, P(C.c_void_p)])
_sig("gps_renderer_destroy", None, [_R])
_sig("gps_renderer_last_error", C.c_char_p, [_R])
_sig("gps_renderer_set_gaussians", C.c_int, [_R, C.c_void_p, u64, P(f32), u32])
_sig("gps_renderer_resize", C.c_int, [_R, u32, u32, u32])
_sig("gps_renderer_reset_accum", C.c_int, [_R])
_sig("gps_renderer_reset_stats", C.c_int, [_R])
_sig("gps_renderer_render_ensembles", C.c_int, [_R, P(GpsCamera), P(RenderParams), u32, u32])
_sig("gps_renderer_render_sorted", C.c_int, [_R, P(GpsCamera), P(RenderParams), u32])
_sig("gps_renderer_probe_demand", C.c_int, [_R, P(GpsCamera), P(RenderParams)])
_sig("gps_renderer_composite", C.c_int, [_R, P(RenderParams), u32, u32, P(u8)])
_sig("gps_renderer_read_accum", C.c_int, [_R, P(f32)])
_sig("gps_renderer_read_stats", C.c_int, [_R, P(GpuStats)])
_sig("gps_renderer_read_sorted", C.c_int, [_R, P(u32), P(u32)])
_sig("gps_renderer_wait_idle", C.c_int, [_R])
_sig("gps_renderer_request_stats", C.c_int, [_R])
_sig("gps_renderer_latest_stats", C.c_int, [_R, P(GpuStats)])
_sig("gps_renderer_stats_issued", u32, [_R])
_sig("gps_renderer_submitted", u32, [_R])
_sig("gps_renderer_accumulated", u32, [_R])
_sig("gps_renderer_info", None, [_R, P(u32), P(u32), P(u32), P(u64), P(u32)])
_sig("gps_estimate_memory", None, [u32, u32, u32, u64, P(u64), P(u64), P(u64)])
_sig("gps_radius_lut", None, [P(f32)])
_sig("gps_renderer_set_debug", C.c_int, [_R, u32])
_sig("gps_renderer_read_counts", C.c_int, [_R, P(u32), P(u32), P(u32), P(u32)])
_sig("gps_renderer_read_depth_id", C.c_int, [_R, P(u64)])
_sig("gps_renderer_read_particle_keys", C.c_int, [_R, P(u64), P(u32), u64, P(u64)])
_sig("gps_renderer_read_particle_counters", C.c_int, [_R, P(u64)])
_sig("gps_sort_pairs_u64", C.c_int, [C.c_int, P(u64), P(u32), u64])
_sig("gps_tuning_get", None, [P(Tuning)])
_sig("gps_tuning_set", C.c_int, [P(Tuning)])
_sig("gps_renderer_autotune", C.c_int, [_R, P(GpsCamera), P(RenderParams), u32, u32, P(Tuning), P(f64)])
_sig("gps_renderer_set_features", C.c_int, [C.c_void_p, C.c_int])
_sig("gps_renderer_set_cull", C.c_int, [C.c_void_p, C.c_int])
_sig("gps_renderer_read_features", C.c_int, [C.c_void_p, P(f32), P(f32)])
_sig("gps_denoiser_param_count", u64, [])
_sig("gps_denoiser_create", C.c_int, [C.c_int, u32, P(C.c_void_p)])
_sig("gps_denoiser_destroy", None, [C.c_void_p])
_sig("gps_denoiser_last_error", C.c_char_p, [C.c_void_p])
_sig("gps_denoiser_set_params", C.c_int, [C.c_void_p, P(f32)])
_sig("gps_denoiser_get_params", C.c_int, [C.c_void_p, P(f32)])
_sig("gps_denoiser_apply", C.c_int, [C.c_void_p, C.c_void_p, f32, P(DenoiseInfo)])
_sig("gps_denoiser_read", C.c_int, [C.c_void_p, P(f32)])
_sig("gps_denoiser_set_fused", C.c_int, [C.c_void_p, C.c_int])
_sig("gps_denoiser_read_features", C.c_int, [C.c_void_p, P(f32), P(f32)])
_sig("gps_denoiser_add_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p])
_sig("gps_denoiser_pairs", u32, [C.c_void_p])
_sig("gps_denoiser_pair_weights", C.c_int, [C.c_void_p, P(f32)])
_sig("gps_denoiser_set_pair_weights", C.c_int, [C.c_void_p, P(f32)])
_sig("gps_denoiser_clear_pairs", C.c_int, [C.c_void_p])
_sig("gps_denoiser_train", C.c_int, [C.c_void_p, u32, u32, f32, u32, P(f32)])
_sig("gps_denoiser_loss_grad", C.c_int, [C.c_void_p, u32, P(f32), P(f32)])
_sig("gps_denoiser_read_output", C.c_int, [C.c_void_p, u32, P(f32)])
_sig("gps_ply_load", C.c_int, [C.c_char_p, u64, P(C.c_void_p)])
_sig("gps_ply_free", None, [C.c_void_p])
_sig("gps_ply_last_error", C.c_char_p, [])
_sig("gps_ply_info", None, [C.c_void_p, P(u64), P(u32), P(C.c_int)])
_sig("gps_ply_read", None, [C.c_void_p, P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)])
_sig("gps_ply_to_gpu", None, [C.c_void_p, C.c_void_p])
_sig("gps_batch_create", C.c_int, [C.c_int, P(C.c_void_p)])
_sig("gps_batch_destroy", None, [C.c_void_p])
_sig("gps_batch_render", C.c_int, [C.c_void_p, P(BatchItem), u32, C.c_int])
_sig("gps_batch_read", C.c_int, [C.c_void_p, u32, P(f32)])
_sig("gps_batch_read_stats", C.c_int, [C.c_void_p, u32, C.c_void_p])
_sig("gps_batch_composite", C.c_int, [C.c_void_p, u32, C.c_void_p, u32, u32, P(C.c_uint8)])
_sig("gps_batch_timing", None, [C.c_void_p, P(f64), P(f64), P(u64)])
_sig("gps_phase0_atomic", C.c_int, [C.c_int, u32, u32, u32, P(u32), P(u32)])
_sig("gps_scan_exclusive_u32", C.c_int, [C.c_int, P(u32), P(u32), u64])
_sig("gps_scan_exclusive_u32_noalloc", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, u64, C.c_void_p])
_sig("gps_math_eval", C.c_int, [C.c_int, C.c_int, P(f32), P(f32), P(f32), u64])
_sig("gps_renderer_profile", C.c_int, [_R, C.c_int])
_sig("gps_renderer_profile_read", C.c_int, [_R, P(KernelTime), C.c_int])

# exported names the header declares (checked by the CPU test suite)
EXPORTED = [
    "gps_render_params_default", "gps_renderer_create", "gps_renderer_destroy", "gps_renderer_last_error",
    "gps_renderer_set_gaussians", "gps_renderer_resize", "gps_renderer_reset_accum", "gps_renderer_reset_stats",
    "gps_renderer_render_ensembles", "gps_renderer_render_sorted", "gps_renderer_probe_demand",
    "gps_renderer_composite", "gps_renderer_read_accum", "gps_renderer_read_stats", "gps_renderer_read_sorted",
    "gps_renderer_request_stats", "gps_renderer_latest_stats", "gps_renderer_stats_issued", "gps_renderer_wait_idle", "gps_renderer_submitted", "gps_renderer_accumulated", "gps_renderer_info",
    "gps_estimate_memory", "gps_radius_lut", "gps_renderer_set_debug", "gps_renderer_read_counts",
    "gps_renderer_read_depth_id", "gps_renderer_read_particle_keys", "gps_renderer_read_particle_counters", "gps_sort_pairs_u64", "gps_phase0_atomic", "gps_scan_exclusive_u32",
    "gps_scan_exclusive_u32_noalloc", "gps_math_eval",
    "gps_renderer_profile", "gps_renderer_profile_read", "gps_tuning_get", "gps_tuning_set",
    "gps_renderer_autotune", "gps_batch_create", "gps_batch_destroy", "gps_batch_render", "gps_batch_read",
    "gps_batch_timing", "gps_renderer_set_features", "gps_renderer_read_features", "gps_renderer_set_cull",
    "gps_denoiser_param_count", "gps_denoiser_create", "gps_denoiser_destroy", "gps_denoiser_last_error",
    "gps_denoiser_set_params", "gps_denoiser_get_params", "gps_denoiser_apply", "gps_denoiser_read",
    "gps_denoiser_read_features", "gps_denoiser_add_pair", "gps_denoiser_pairs", "gps_denoiser_clear_pairs",
    "gps_denoiser_train", "gps_denoiser_loss_grad", "gps_denoiser_read_output", "gps_denoiser_pair_weights",
    "gps_denoiser_set_pair_weights", "gps_denoiser_set_fused", "gps_ply_load", "gps_ply_free", "gps_ply_last_error", "gps_ply_info",
    "gps_ply_read", "gps_ply_to_gpu", "gps_batch_read_stats", "gps_batch_composite",
]

DEBUG_KEEP_BUFFERS = 1


def _ptr(a, t):
    return a.ctypes.data_as(P(t))


def radius_lut():
    """crates/gps-core/src/lib.rs:624-636, computed by the library (f64, rounded to f32)."""
    t = np.zeros(65 * 129, np.float32)
    lib.gps_radius_lut(_ptr(t, f32))
    return t


def estimate_memory(w, h, spp_side, n):
    a, b, c = u64(), u64(), u64()
    lib.gps_estimate_memory(w, h, spp_side, n, C.byref(a), C.byref(b), C.byref(c))
    return dict(depth_winner_bytes=a.value, accum_bytes=b.value, splat_bytes=c.value)


class GpsRenderer:
    """gps.rs:302-1065 on one HIP device."""

    def __init__(self, width, height, spp_side, device=0):
        h = C.c_void_p()
        if lib.gps_renderer_create(width, height, spp_side, device, C.byref(h)) != 0:
            raise GpsError(lib.gps_renderer_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.gps_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise GpsError(lib.gps_renderer_last_error(self._h).decode())

    # ---- scene / buffers
    def set_gaussians(self, data):
        return self.set_gaussians_sh(data, None, 0)

    def set_gaussians_sh(self, data, sh_rest, degree):
        data = np.ascontiguousarray(data)
        if data.dtype != GAUSSIAN_DTYPE:
            data = np.ascontiguousarray(data, dtype=np.float32).reshape(-1, 16).view(GAUSSIAN_DTYPE).reshape(-1)
        n = data.shape[0]
        sh = None
        if sh_rest is not None and degree >= 1 and n > 0:
            stride = (min(degree, 3) + 1) ** 2 - 1
            sh = np.ascontiguousarray(sh_rest, np.float32).reshape(-1)
            if sh.size != n * 3 * stride:
                raise GpsError(f"SH data has {sh.size} values, expected {n * 3 * stride} for {n} splats at degree "
                               f"{degree}")
        self._check(lib.gps_renderer_set_gaussians(self._h, data.ctypes.data if n else None, n,
                                                   _ptr(sh, f32) if sh is not None else None,
                                                   degree if sh is not None else 0))

    def resize(self, width, height, spp_side):
        self._check(lib.gps_renderer_resize(self._h, width, height, spp_side))

    def reset_accum(self):
        self._check(lib.gps_renderer_reset_accum(self._h))

    def reset_stats(self):
        self._check(lib.gps_renderer_reset_stats(self._h))

    # ---- rendering
    def render_ensembles(self, cam, params, first_seed, count):
        self._check(lib.gps_renderer_render_ensembles(self._h, C.byref(cam), C.byref(params), first_seed & 0xFFFFFFFF,
                                                      count))

    def render_sorted(self, cam, params, as_ensembles=1):
        self._check(lib.gps_renderer_render_sorted(self._h, C.byref(cam), C.byref(params), as_ensembles))

    def probe_demand(self, cam, params):
        self._check(lib.gps_renderer_probe_demand(self._h, C.byref(cam), C.byref(params)))

    def composite(self, params, out_w=None, out_h=None):
        w, h = self.size()
        ow, oh = out_w or w, out_h or h
        out = np.zeros((oh, ow, 4), np.uint8)
        self._check(lib.gps_renderer_composite(self._h, C.byref(params), ow, oh, _ptr(out, u8)))
        return out

    # ---- readbacks
    def read_accum(self):
        w, h = self.size()
        out = np.zeros((h, w, 4), np.float32)
        self._check(lib.gps_renderer_read_accum(self._h, _ptr(out, f32)))
        return out

    def set_cull(self, enable=True):
        """Skip PBVR splats whose particles all miss the image (same image; fewer counted points)."""
        self._check(lib.gps_renderer_set_cull(self._h, 1 if enable else 0))

    def set_features(self, enable=True):
        """Accumulate the denoiser's per-pixel features (sums of squares, covered subpixels, nearest depth)."""
        self._check(lib.gps_renderer_set_features(self._h, 1 if enable else 0))

    def read_features(self):
        """(moments (h, w, 4): sum of squared ensemble values r, g, b and covered subpixels; zmin (h, w))."""
        w, h = self.size()
        mom = np.zeros((h, w, 4), np.float32)
        zmin = np.zeros((h, w), np.float32)
        self._check(lib.gps_renderer_read_features(self._h, _ptr(mom, f32), _ptr(zmin, f32)))
        return mom, zmin

    def read_stats(self):
        s = GpuStats()
        self._check(lib.gps_renderer_read_stats(self._h, C.byref(s)))
        return s

    def request_stats(self):
        """gps.rs:1005-1027: non-blocking counter readback (no-op while one is in flight)."""
        self._check(lib.gps_renderer_request_stats(self._h))

    def latest_stats(self):
        """gps.rs:1035-1038: the last delivered readback, its issue number in .issue."""
        s = GpuStats()
        self._check(lib.gps_renderer_latest_stats(self._h, C.byref(s)))
        return s

    def stats_issued(self):
        """gps.rs:1029-1033"""
        return int(lib.gps_renderer_stats_issued(self._h))

    def read_sorted(self):
        n = self.gaussian_count()
        k = np.zeros(max(n, 1), np.uint32)
        v = np.zeros(max(n, 1), np.uint32)
        self._check(lib.gps_renderer_read_sorted(self._h, _ptr(k, u32), _ptr(v, u32)))
        return k[:n], v[:n]

    def wait_idle(self):
        self._check(lib.gps_renderer_wait_idle(self._h))

    def submitted(self):
        return lib.gps_renderer_submitted(self._h)

    def accumulated(self):
        return lib.gps_renderer_accumulated(self._h)

    def _info(self):
        w, h, s, d = u32(), u32(), u32(), u32()
        n = u64()
        lib.gps_renderer_info(self._h, C.byref(w), C.byref(h), C.byref(s), C.byref(n), C.byref(d))
        return w.value, h.value, s.value, n.value, d.value

    def size(self):
        w, h, _, _, _ = self._info()
        return w, h

    def spp_side(self):
        return self._info()[2]

    def gaussian_count(self):
        return self._info()[3]

    def sh_max_degree(self):
        return self._info()[4]

    # ---- bit-exact test hooks
    def set_debug(self, flags):
        self._check(lib.gps_renderer_set_debug(self._h, flags))

    def read_counts(self):
        n = self.gaussian_count()
        c = np.zeros(max(n, 1), np.uint32)
        o = np.zeros(max(n, 1), np.uint32)
        t, sk = u32(), u32()
        self._check(lib.gps_renderer_read_counts(self._h, _ptr(c, u32), _ptr(o, u32), C.byref(t), C.byref(sk)))
        return dict(counts=c[:n], offsets=o[:n], total=t.value, skipped=sk.value)

    def read_depth_id(self):
        w, h, s, _, _ = self._info()
        out = np.zeros(w * h * s * s, np.uint64)
        self._check(lib.gps_renderer_read_depth_id(self._h, _ptr(out, u64)))
        return out

    def read_particle_counters(self):
        """Particle kernel counters since the last reset_stats: particles that reached a subpixel (hits), and the
        count-only LDS buffers' high-water marks with their capacities under the current tuning."""
        out = (u64 * 5)()
        self._check(lib.gps_renderer_read_particle_counters(self._h, out))
        return dict(hits=int(out[0]), queue_peak=int(out[1]), queue_cap=int(out[2]), work_peak=int(out[3]),
                    work_cap=int(out[4]))

    def read_queue_peaks(self):
        """(queue peak, queue capacity, work peak, work capacity) of read_particle_counters."""
        c = self.read_particle_counters()
        return c["queue_peak"], c["queue_cap"], c["work_peak"], c["work_cap"]

    def read_particle_keys(self, max_count=None):
        n = u64()
        lib.gps_renderer_read_particle_keys(self._h, None, None, 0, C.byref(n))
        m = n.value if max_count is None else min(n.value, max_count)
        k = np.zeros(max(m, 1), np.uint64)
        s = np.zeros(max(m, 1), np.uint32)
        self._check(lib.gps_renderer_read_particle_keys(self._h, _ptr(k, u64), _ptr(s, u32), m, C.byref(n)))
        return k[:m], s[:m]

    def autotune(self, cam, params, first_seed=1_000_000, trial_ensembles=2):
        t = Tuning()
        ms = f64()
        self._check(lib.gps_renderer_autotune(self._h, C.byref(cam), C.byref(params), first_seed, trial_ensembles,
                                              C.byref(t), C.byref(ms)))
        return t.as_dict(), ms.value

    # ---- timing
    def profile(self, enable=True):
        self._check(lib.gps_renderer_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        buf = (KernelTime * 64)()
        n = lib.gps_renderer_profile_read(self._h, buf, 64)
        if n < 0:
            raise GpsError("profile read failed")
        return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(min(n, 64))}


BATCH_GRAPH, BATCH_PER_ITEM = 1, 2


class Batch:
    """Many (renderer, view, ensembles) items through one set of fused kernels (or each item's own launches with
    per_item), captured into one hipGraph with use_graph."""

    def __init__(self, device=0):
        h = C.c_void_p()
        if lib.gps_batch_create(device, C.byref(h)) != 0:
            raise GpsError(lib.gps_renderer_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.gps_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, items, use_graph=True, per_item=False):
        """items: list of (GpsRenderer, GpsCamera, RenderParams, first_seed, count)."""
        self._keep = [(cam, p) for _, cam, p, _, _ in items]
        arr = (BatchItem * len(items))()
        for i, (r, cam, p, seed, count) in enumerate(items):
            arr[i].renderer = r._h
            arr[i].camera = C.addressof(cam)
            arr[i].params = C.addressof(p)
            arr[i].first_seed = seed & 0xFFFFFFFF
            arr[i].count = count
        self._sizes = [r.size() for r, _, _, _, _ in items]
        flags = (BATCH_GRAPH if use_graph else 0) | (BATCH_PER_ITEM if per_item else 0)
        if lib.gps_batch_render(self._h, arr, len(items), flags) != 0:
            raise GpsError(lib.gps_renderer_last_error(None).decode())

    def read(self, i):
        w, h = self._sizes[i]
        out = np.zeros((h, w, 4), np.float32)
        if lib.gps_batch_read(self._h, i, _ptr(out, f32)) != 0:
            raise GpsError("batch read failed")
        return out

    def composite(self, i, params, out_w=None, out_h=None):
        w, h = self._sizes[i]
        ow, oh = out_w or w, out_h or h
        out = np.zeros((oh, ow, 4), np.uint8)
        if lib.gps_batch_composite(self._h, i, C.byref(params), ow, oh, out.ctypes.data_as(P(C.c_uint8))) != 0:
            raise GpsError("batch composite failed")
        return out

    def read_stats(self, i):
        s = GpuStats()
        if lib.gps_batch_read_stats(self._h, i, C.byref(s)) != 0:
            raise GpsError("batch stats read failed")
        return s

    def timing(self):
        g, rec, nodes = f64(), f64(), u64()
        lib.gps_batch_timing(self._h, C.byref(g), C.byref(rec), C.byref(nodes))
        return dict(gpu_ms=g.value, record_ms=rec.value, graph_nodes=nodes.value)


# ------------------------------------------------------------------ 3DGS PLY ingest (crates/gps-io/src/lib.rs)
PLY_ERRORS = {1: "NotPly", 2: "BadHeader", 3: "UnsupportedFormat", 4: "MissingProperty", 5: "TooManyVertices",
              6: "Truncated", 7: "BadAscii", 8: "NonFinite", 9: "BadShCount"}
PLY_MAX_VERTICES = 200_000_000


class PlyError(GpsError):
    """gps_io::PlyError: `kind` is the variant name, str(e) its message."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code
        self.kind = PLY_ERRORS.get(code, "Unknown")


class PlyGaussians:
    """gps_io::Gaussians: file-form struct-of-arrays (log-scales, opacity logits, (w, x, y, z) rotations,
    f_dc, channel-major f_rest per splat)."""

    def __init__(self, handle):
        c, d, b = u64(), u32(), C.c_int()
        lib.gps_ply_info(handle, C.byref(c), C.byref(d), C.byref(b))
        self.count, self.sh_degree, self.binary = int(c.value), int(d.value), bool(b.value)
        n, r = self.count, self.sh_rest_per_channel()
        self.position = np.zeros((n, 3), np.float32)
        self.log_scale = np.zeros((n, 3), np.float32)
        self.rot_wxyz = np.zeros((n, 4), np.float32)
        self.opacity_logit = np.zeros(n, np.float32)
        self.f_dc = np.zeros((n, 3), np.float32)
        self.sh_rest = np.zeros(n * 3 * r, np.float32)
        lib.gps_ply_read(handle, _ptr(self.position, f32), _ptr(self.log_scale, f32), _ptr(self.rot_wxyz, f32),
                         _ptr(self.opacity_logit, f32), _ptr(self.f_dc, f32),
                         _ptr(self.sh_rest, f32) if self.sh_rest.size else None)
        self._gpu = np.zeros(n, GAUSSIAN_DTYPE)
        if n:
            lib.gps_ply_to_gpu(handle, self._gpu.ctypes.data)

    def sh_rest_per_channel(self):
        return (self.sh_degree + 1) ** 2 - 1 if self.sh_degree else 0

    def to_gpu(self):
        """GpuGaussian::from_activated(activated(i)) for every splat (gps.rs:41-49)."""
        return self._gpu.copy()


def load_ply(data):
    """gps_io::load_ply: `data` is the file's bytes or a path."""
    if not isinstance(data, (bytes, bytearray, memoryview)):
        with open(data, "rb") as fh:
            data = fh.read()
    data = bytes(data)
    h = C.c_void_p()
    rc = lib.gps_ply_load(data, len(data), C.byref(h))
    if rc != 0:
        raise PlyError(rc, lib.gps_ply_last_error().decode())
    try:
        return PlyGaussians(h)
    finally:
        lib.gps_ply_free(h)


def robust_sphere(points):
    """OrbitCamera::robust_sphere (crates/gps-core/src/camera.rs:70-84): per-axis median centre and the 75th
    percentile distance to it; None for no points."""
    p = np.asarray(points, np.float64).reshape(-1, 3)
    if p.shape[0] == 0:
        return None
    c = np.array([np.sort(p[:, a])[p.shape[0] // 2] for a in range(3)])
    d = np.sort(np.sqrt(((p - c) ** 2).sum(axis=1)))
    return c, float(d[min(int(d.size * 0.75), d.size - 1)])


def frame_sphere_distance(radius, fov_y=0.9):
    """OrbitCamera::frame_sphere (camera.rs:87-90): the orbit distance that fills about 80% of the view."""
    return (max(radius, 1e-6) / math.sin(0.5 * fov_y)) * 1.25


def denoiser_param_count():
    return int(lib.gps_denoiser_param_count())


def _denoiser_layers():
    """(cin as stored, cout) per layer: 8 -> 32, then 32 -> 32 layers, then 32 -> 3; the library's parameter count
    fixes how many hidden layers it was built with."""
    first, hidden, last = 32 * 9 * 8 + 32, 32 * 9 * 32 + 32, 3 * 9 * 32 + 3
    k, rem = divmod(denoiser_param_count() - first - last, hidden)
    if rem or k < 1:
        raise GpsError(f"unexpected denoiser parameter count {denoiser_param_count()}")
    return ((8, 32),) + ((32, 32),) * k + ((32, 3),)


DENOISER_LAYERS = _denoiser_layers()


def denoiser_split(params):
    """Parameter vector -> [(W (cout, 3, 3, cin), b (cout,))] per layer."""
    out, o = [], 0
    for cin, cout in DENOISER_LAYERS:
        w = np.asarray(params[o:o + cout * 9 * cin]).reshape(cout, 3, 3, cin)
        o += cout * 9 * cin
        out.append((w, np.asarray(params[o:o + cout])))
        o += cout
    return out


class Denoiser:
    """Learned denoiser (5 layers of 3x3 convolutions, 32 channels, BF16 on MFMA) of a renderer's accumulation."""

    def __init__(self, seed=1, device=0):
        h = C.c_void_p()
        if lib.gps_denoiser_create(device, seed, C.byref(h)) != 0:
            raise GpsError("denoiser creation failed")
        self._h = h
        self._size = None

    def close(self):
        if getattr(self, "_h", None):
            lib.gps_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise GpsError(lib.gps_denoiser_last_error(self._h).decode())
        return rc

    def get_params(self):
        out = np.zeros(denoiser_param_count(), np.float32)
        self._check(lib.gps_denoiser_get_params(self._h, _ptr(out, f32)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == denoiser_param_count()
        self._check(lib.gps_denoiser_set_params(self._h, _ptr(p, f32)))

    def save(self, path):
        self.get_params().tofile(path)

    def load(self, path):
        self.set_params(np.fromfile(path, np.float32))

    def apply(self, renderer, fallback_ratio=0.0):
        info = DenoiseInfo()
        self._check(lib.gps_denoiser_apply(self._h, renderer._h, fallback_ratio, C.byref(info)))
        self._size = (info.width, info.height)
        return dict(gpu_ms=info.gpu_ms, tiles=info.tiles, fallback_tiles=info.fallback_tiles,
                    fallback_rate=info.fallback_tiles / max(info.tiles, 1))

    def set_fused(self, enable=True):
        """apply's forward: the fused five-layer kernel (default) or the layer-by-layer kernels (the same bits).
        Returns False when this build has no fused kernel."""
        return self._check(lib.gps_denoiser_set_fused(self._h, 1 if enable else 0)) == 0

    def read(self):
        w, h = self._size
        out = np.zeros((h, w, 4), np.float32)
        self._check(lib.gps_denoiser_read(self._h, _ptr(out, f32)))
        return out

    def read_features(self):
        w, h = self._size
        x8 = np.zeros((h, w, 8), np.float32)
        mean = np.zeros((h, w, 4), np.float32)
        self._check(lib.gps_denoiser_read_features(self._h, _ptr(x8, f32), _ptr(mean, f32)))
        return x8, mean

    def add_pair(self, inp, target):
        return self._check(lib.gps_denoiser_add_pair(self._h, inp._h, target._h))

    def pairs(self):
        return int(lib.gps_denoiser_pairs(self._h))

    def pair_weights(self):
        out = np.zeros(max(self.pairs(), 1), np.float32)
        self._check(lib.gps_denoiser_pair_weights(self._h, _ptr(out, f32)))
        return out[:self.pairs()]

    def set_pair_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.pairs()
        self._check(lib.gps_denoiser_set_pair_weights(self._h, _ptr(w, f32)))

    def clear_pairs(self):
        self._check(lib.gps_denoiser_clear_pairs(self._h))

    def train(self, steps, batch, lr=1e-3, seed=1):
        losses = np.zeros(steps, np.float32)
        self._check(lib.gps_denoiser_train(self._h, steps, batch, lr, seed, _ptr(losses, f32)))
        return losses

    def loss_grad(self, batch):
        loss = np.zeros(1, np.float32)
        grad = np.zeros(denoiser_param_count(), np.float32)
        self._check(lib.gps_denoiser_loss_grad(self._h, batch, _ptr(loss, f32), _ptr(grad, f32)))
        return float(loss[0]), grad

    def read_output(self, images, w, h):
        out = np.zeros((images, h, w, 4), np.float32)
        self._check(lib.gps_denoiser_read_output(self._h, images, _ptr(out, f32)))
        return out


# ------------------------------------------------------------------ unit entry points
def phase0_atomic(seed, pixels, samples, device=0):
    d = np.zeros(pixels, np.uint32)
    c = np.zeros(pixels, np.uint32)
    if lib.gps_phase0_atomic(device, seed, pixels, samples, _ptr(d, u32), _ptr(c, u32)) != 0:
        raise GpsError("phase-0 atomic kernel failed")
    return d, c


# gps_scan_exclusive_u32_noalloc's limit and its error code (include/gps_hip.h)
SCAN_NOALLOC_MAX_N = 1 << 28
SCAN_NOALLOC_TOO_LARGE = -2


def scan_exclusive_u32(x, device=0):
    x = np.ascontiguousarray(x, np.uint32)
    out = np.zeros_like(x)
    if lib.gps_scan_exclusive_u32(device, _ptr(x, u32), _ptr(out, u32), x.size) != 0:
        raise GpsError("scan kernel failed")
    return out


def get_tuning():
    t = Tuning()
    lib.gps_tuning_get(C.byref(t))
    return t.as_dict()


def set_tuning(**kw):
    t = Tuning()
    lib.gps_tuning_get(C.byref(t))
    for k, v in kw.items():
        setattr(t, k, v)
    if lib.gps_tuning_set(C.byref(t)) != 0:
        raise GpsError(f"invalid tuning {t.as_dict()}")
    return t.as_dict()


def sort_pairs_u64(keys, vals, device=0):
    """Stable LSD radix sort on the GPU (returns sorted copies)."""
    k = np.array(keys, np.uint64, copy=True)
    v = np.array(vals, np.uint32, copy=True)
    if lib.gps_sort_pairs_u64(device, _ptr(k, u64), _ptr(v, u32), k.size) != 0:
        raise GpsError("sort failed")
    return k, v


MATH_FN = {"exp": 0, "log": 1, "sin": 2, "cos": 3, "sqrt": 4, "div": 5}


def math_eval(fn, x, y=None, device=0):
    x = np.ascontiguousarray(x, np.float32)
    yy = np.ascontiguousarray(y, np.float32) if y is not None else None
    out = np.zeros_like(x)
    if lib.gps_math_eval(device, MATH_FN[fn], _ptr(x, f32), _ptr(yy, f32) if yy is not None else None,
                         _ptr(out, f32), x.size) != 0:
        raise GpsError("math kernel failed")
    return out


# ------------------------------------------------------------------ cameras (oracle.rs:53-70, camera.rs:31-40)
def _normalize(v):
    v = np.asarray(v, np.float64)
    return v * (1.0 / math.sqrt(float(v @ v)))


def look_at(eye, target, up, fov_y, w, h):
    """OracleCamera::look_at: +Z forward, image +Y down; rows of view_rot are (right, down, forward)."""
    eye = np.asarray(eye, np.float64)
    f = _normalize(np.asarray(target, np.float64) - eye)
    right = _normalize(np.cross(f, np.asarray(up, np.float64)))
    down = np.cross(f, right)
    c = GpsCamera()
    c.view_rot[:] = [*right, *down, *f]
    c.view_pos[:] = list(eye)
    focal = 0.5 * h / math.tan(0.5 * fov_y)
    c.focal_x = c.focal_y = focal
    c.cx, c.cy = 0.5 * w, 0.5 * h
    c.low_pass = 0.3
    return c


def orbit_eye(target, distance, yaw, pitch, up):
    """OrbitCamera::eye, camera.rs:31-40."""
    up = _normalize(up)
    helper = np.array([1.0, 0.0, 0.0]) if abs(up[0]) < 0.9 else np.array([0.0, 0.0, 1.0])
    a = _normalize(np.cross(up, helper))
    b = np.cross(up, a)
    horiz = a * math.cos(yaw) + b * math.sin(yaw)
    d = horiz * math.cos(pitch) + up * math.sin(pitch)
    return np.asarray(target, np.float64) + d * distance


def orbit_camera(target, distance, yaw, pitch, up, fov_y, w, h):
    return look_at(orbit_eye(target, distance, yaw, pitch, up), target, up, fov_y, w, h)


def default_camera(focal=200.0, low_pass=0.3):
    """OracleCamera::default() (oracle.rs:44-48) with a chosen focal length."""
    c = GpsCamera()
    c.view_rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    c.view_pos[:] = [0, 0, 0]
    c.focal_x = c.focal_y = focal
    c.cx = c.cy = 0.0
    c.low_pass = low_pass
    return c
