"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It loads oracle/liboracle.so, built by oracle/Makefile:

* ``orc_*``: f64 restatement of crates/gps-core (lib.rs, oracle.rs, camera.rs)
* ``mir_*``: f32 restatement of shaders/gps_splat.wgsl and the orchestration
  of crates/gps-render/src/gps.rs:812-925
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()

u32, u64, f32, f64, sz = C.c_uint32, C.c_uint64, C.c_float, C.c_double, C.c_size_t
P = C.POINTER


class OrcGaussian(C.Structure):
    """oracle.rs:14-23 (f64; rot in glam (x, y, z, w) order)"""
    _fields_ = [("pos", f64 * 3), ("log_scale", f64 * 3), ("rot_xyzw", f64 * 4), ("opacity", f64), ("color", f64 * 3)]


class OrcCamera(C.Structure):
    """oracle.rs:33-42; view_rot row-major"""
    _fields_ = [("view_rot", f64 * 9), ("view_pos", f64 * 3), ("focal_x", f64), ("focal_y", f64), ("cx", f64),
                ("cy", f64), ("low_pass", f64)]


class OrcStats(C.Structure):
    _fields_ = [("points", f64), ("candidates", f64), ("subpixel_writes", f64)]


class OrcMcOptions(C.Structure):
    """oracle.rs:204-226"""
    _fields_ = [("footprint_subpixels", u32), ("adaptive_kappa", f64), ("footprint_max", u32), ("compensate_blur", u32),
                ("density_scale", f64), ("per_particle_seed", u32), ("max_points_per_splat", u32)]


class OrcParticleOptions(C.Structure):
    """oracle.rs:316-350"""
    _fields_ = [("level", u32), ("rule", u32), ("radial_correction", u32), ("base_k", f64),
                ("reference_pixel_length", f64), ("near_z", f64), ("low_pass_jitter", u32), ("centre_depth", u32),
                ("linearized_projection", u32), ("density_scale", f64), ("per_particle_seed", u32)]


class OrcCalibInputs(C.Structure):
    _fields_ = [(n, f64) for n in ("object_depth", "splat_depth", "focal_x", "focal_y", "near_z",
                                   "reference_pixel_length", "footprint_area", "spp")]


class MGaussian(C.Structure):
    """GpuGaussian, gps.rs:29-38"""
    _fields_ = [("pos_opacity", f32 * 4), ("log_scale", f32 * 4), ("rot_wxyz", f32 * 4), ("color", f32 * 4)]


class MPbvr(C.Structure):
    _fields_ = [("calibration", u32), ("radial_correction", u32), ("centre_depth", u32), ("low_pass_jitter", u32),
                ("base_k", f32), ("reference_pixel_length", f32)]


class MParams(C.Structure):
    """RenderParams, gps.rs:189-229"""
    _fields_ = [("spp_side", u32), ("density_scale", f32), ("max_points_per_splat", u32), ("near", f32),
                ("background", f32 * 3), ("exposure", f32), ("srgb_output", u32), ("exact_radius", u32),
                ("sh_degree", u32), ("path", u32), ("method", u32), ("pbvr", MPbvr)]


def default_mc_options(**kw):
    o = OrcMcOptions(footprint_subpixels=1, adaptive_kappa=0.0, footprint_max=4, compensate_blur=0, density_scale=1.0,
                     per_particle_seed=0, max_points_per_splat=0xFFFFFFFF)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_particle_options(**kw):
    o = OrcParticleOptions(level=3, rule=1, radial_correction=0, base_k=512.0, reference_pixel_length=0.01,
                           near_z=0.05, low_pass_jitter=1, centre_depth=0, linearized_projection=0, density_scale=1.0,
                           per_particle_seed=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _sig(name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


_sig("orc_pcg_hash", u32, [u32])
_sig("orc_particle_seed", u32, [u32, u32, u32, u32])
_sig("orc_dilog", f64, [f64])
_sig("orc_inv_dilog", f64, [f64])
_sig("orc_sigmoid", f64, [f64])
_sig("orc_expected_point_count", f64, [f64, f64])
_sig("orc_ordered_depth_key", u32, [f32])
_sig("orc_ordered_depth_key_inverse", f32, [u32])
_sig("orc_sample_corrected_radius", f64, [f64, f64])
_sig("orc_radial_keep_probability", f64, [f64, f64])
_sig("orc_expected_coverage", f64, [f64, f64])
_sig("orc_poisson_mean", f64, [f64, u32, u32])
_sig("orc_calibrated_count", f64, [u32, u32, f64, f64, f64, P(OrcCalibInputs)])
_sig("orc_radius_lut", None, [P(f32)])
_sig("orc_radius_from_lut", f64, [P(f32), f64, f64])
_sig("orc_eval_sh", None, [u32, P(f64), P(f64), P(f64), u32, P(f64)])
_sig("orc_look_at", None, [P(f64), P(f64), P(f64), f64, u32, u32, P(OrcCamera)])
_sig("orc_orbit_camera", None, [P(f64), f64, f64, f64, P(f64), f64, u32, u32, P(OrcCamera)])
_sig("orc_render_analytic", None, [P(OrcGaussian), sz, P(OrcCamera), sz, sz, P(f64), P(f64)])
_sig("orc_render_analytic_area", None, [P(OrcGaussian), sz, P(OrcCamera), sz, sz, P(f64), u32, P(f64)])
_sig("orc_render_monte_carlo", None, [P(OrcGaussian), sz, P(OrcCamera), sz, sz, sz, sz, P(f64), u32,
                                      P(OrcMcOptions), P(f64), P(OrcStats)])
_sig("orc_render_particles_3d", None, [P(OrcGaussian), sz, P(OrcCamera), sz, sz, sz, sz, P(f64), u32,
                                       P(OrcParticleOptions), P(f64), P(OrcStats)])
_sig("orc_psnr", f64, [P(f64), P(f64), sz])
_sig("orc_ssim", f64, [P(f64), P(f64), sz, sz])

_sig("mir_pcg", u32, [u32])
_sig("mir_particle_seed", u32, [u32, u32, u32])
_sig("mir_dilog", f32, [f32])
_sig("mir_inv_dilog", f32, [f32, f32])
_sig("mir_corrected_radius", f32, [f32, f32])
_sig("mir_corrected_radius_lut", f32, [P(f32), f32, f32])
_sig("mir_radial_keep", f32, [f32, f32])
_sig("mir_poisson", u32, [f32, u32])
for _n in ("mir_expf_n", "mir_logf_n", "mir_sinf_n", "mir_cosf_n"):
    _sig(_n, None, [P(f32), P(f32), sz])
_sig("mir_set_features", None, [P(f32), P(f32)])
_sig("mir_set_threads", None, [C.c_int])
_sig("mir_render_ensembles", C.c_int, [P(MGaussian), u32, P(f32), u32, P(f32), P(OrcCamera), P(MParams), u32, u32,
                                        u32, u32, u32, P(f32), P(u64), P(u32), P(u32), P(u64), P(u32), P(u32)])
_sig("mir_splat_colours", None, [P(MGaussian), u32, P(f32), u32, P(OrcCamera), P(MParams), u32, u32, P(f32)])
_sig("mir_render_sorted", C.c_int, [P(MGaussian), u32, P(f32), u32, P(OrcCamera), P(MParams), u32, u32, u32, P(f32),
                                     P(u64), P(u32), P(u32), P(u32)])
_sig("mir_composite", None, [P(f32), u32, u32, f32, u32, u32, u32, P(C.c_uint8)])
_sig("mir_phase0", None, [u32, u32, u32, u32, P(u32), P(u32)])


def ptr(a, t):
    return a.ctypes.data_as(P(t))


# ------------------------------------------------------------------ helpers
def camera_array(cam):
    """OrcCamera from anything with view_rot (3x3), view_pos, focal_x, focal_y, cx, cy, low_pass."""
    if isinstance(cam, OrcCamera):
        return cam
    c = OrcCamera()
    c.view_rot[:] = [float(v) for v in np.asarray(cam["view_rot"], dtype=np.float64).reshape(9)]
    c.view_pos[:] = [float(v) for v in cam["view_pos"]]
    for k in ("focal_x", "focal_y", "cx", "cy", "low_pass"):
        setattr(c, k, float(cam[k]))
    return c


def look_at(eye, target, up, fov_y, w, h):
    c = OrcCamera()
    e = (f64 * 3)(*eye)
    t = (f64 * 3)(*target)
    u = (f64 * 3)(*up)
    lib.orc_look_at(e, t, u, fov_y, w, h, C.byref(c))
    return c


def orbit_camera(target, distance, yaw, pitch, up, fov_y, w, h):
    c = OrcCamera()
    lib.orc_orbit_camera((f64 * 3)(*target), distance, yaw, pitch, (f64 * 3)(*up), fov_y, w, h, C.byref(c))
    return c


def default_camera(focal=200.0):
    """OracleCamera::default() with a focal length (oracle.rs:44-48)."""
    c = OrcCamera()
    c.view_rot[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    c.view_pos[:] = [0, 0, 0]
    c.focal_x = c.focal_y = focal
    c.cx = c.cy = 0.0
    c.low_pass = 0.3
    return c


def gaussians_f64(lst):
    arr = (OrcGaussian * max(len(lst), 1))()
    for i, g in enumerate(lst):
        arr[i].pos[:] = g["pos"]
        arr[i].log_scale[:] = g["log_scale"]
        arr[i].rot_xyzw[:] = g.get("rot_xyzw", (0.0, 0.0, 0.0, 1.0))
        arr[i].opacity = g["opacity"]
        arr[i].color[:] = g["color"]
    return arr


def render_analytic(scene, cam, w, h, bg):
    out = np.zeros((h, w, 3), np.float64)
    arr = gaussians_f64(scene)
    lib.orc_render_analytic(arr, len(scene), C.byref(cam), w, h, (f64 * 3)(*bg), ptr(out, f64))
    return out


def render_analytic_area(scene, cam, w, h, bg, ss=8):
    out = np.zeros((h, w, 3), np.float64)
    arr = gaussians_f64(scene)
    lib.orc_render_analytic_area(arr, len(scene), C.byref(cam), w, h, (f64 * 3)(*bg), ss, ptr(out, f64))
    return out


def render_monte_carlo(scene, cam, w, h, spp_side, sets, bg, seed, opt=None):
    out = np.zeros((h, w, 3), np.float64)
    st = OrcStats()
    arr = gaussians_f64(scene)
    opt = opt or default_mc_options()
    lib.orc_render_monte_carlo(arr, len(scene), C.byref(cam), w, h, spp_side, sets, (f64 * 3)(*bg), seed,
                               C.byref(opt), ptr(out, f64), C.byref(st))
    return out, st


def render_particles_3d(scene, cam, w, h, spp_side, sets, bg, seed, opt=None):
    """scene: a list of dicts, or an (n, 14) f64 array of OrcGaussian rows (pos, log_scale, rot xyzw, opacity,
    colour)."""
    out = np.zeros((h, w, 3), np.float64)
    st = OrcStats()
    if isinstance(scene, np.ndarray):
        keep = np.ascontiguousarray(scene, np.float64).reshape(-1, 14)
        assert C.sizeof(OrcGaussian) == 14 * 8
        arr = keep.ctypes.data_as(P(OrcGaussian))
    else:
        arr = gaussians_f64(scene)
    opt = opt or default_particle_options()
    lib.orc_render_particles_3d(arr, len(scene), C.byref(cam), w, h, spp_side, sets, (f64 * 3)(*bg), seed,
                                C.byref(opt), ptr(out, f64), C.byref(st))
    return out, st


def psnr(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib.orc_psnr(ptr(a, f64), ptr(b, f64), a.size // 3)


def ssim(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    h, w = a.shape[0], a.shape[1]
    return lib.orc_ssim(ptr(a, f64), ptr(b, f64), w, h)


def radius_lut():
    t = np.zeros(65 * 129, np.float32)
    lib.orc_radius_lut(ptr(t, f32))
    return t


_LUT = None


def cached_lut():
    global _LUT
    if _LUT is None:
        _LUT = radius_lut()
    return _LUT


STAT_NAMES = ("points", "truncated_splats", "dropped_points", "visible_splats", "orphan_subpixels", "candidates",
              "undercovered_splats", "skipped_ensembles", "ensembles", "probe_demand")


def mirror_render(splats, sh_rest, sh_degree, cam, params, w, h, first_seed, count, accum=None, stats=None,
                  want_buffers=False, lut=None, features=None):
    """f32 mirror of GpsRenderer::render_ensembles.  splats: structured ndarray / bytes of GpuGaussian (64 B each).
    features: optional (moments (h, w, 4) f32, zmin (h, w) f32) updated in place like the renderer's
    denoiser features."""
    sp = np.ascontiguousarray(splats).view(np.float32).reshape(-1, 16)
    n = sp.shape[0]
    spp_side = params.spp_side
    if accum is None:
        accum = np.zeros((h, w, 4), np.float32)
    if stats is None:
        stats = np.zeros(16, np.uint64)
    sh = np.ascontiguousarray(sh_rest if sh_rest is not None else np.zeros(1, np.float32), np.float32)
    lut = cached_lut() if lut is None else lut
    out = {}
    counts = offsets = depth_id = None
    total = u32(0)
    skipped = u32(0)
    if want_buffers:
        counts = np.zeros(max(n, 1), np.uint32)
        offsets = np.zeros(max(n, 1), np.uint32)
        depth_id = np.zeros(w * h * spp_side * spp_side, np.uint64)
    if features is not None:
        assert features[0].dtype == np.float32 and features[0].size == 4 * w * h and features[0].flags.c_contiguous
        assert features[1].dtype == np.float32 and features[1].size == w * h and features[1].flags.c_contiguous
        lib.mir_set_features(ptr(features[0], f32), ptr(features[1], f32))
    rc = lib.mir_render_ensembles(sp.ctypes.data_as(P(MGaussian)), n, ptr(sh, f32), sh_degree if sh_rest is not None else 0,
                                  ptr(lut, f32), C.byref(camera_array(cam)), C.byref(params), w, h, spp_side,
                                  first_seed, count, ptr(accum, f32), ptr(stats, u64),
                                  ptr(counts, u32) if want_buffers else None,
                                  ptr(offsets, u32) if want_buffers else None,
                                  ptr(depth_id, u64) if want_buffers else None, C.byref(total), C.byref(skipped))
    if features is not None:
        lib.mir_set_features(None, None)
    if rc != 0:
        raise ValueError("mirror render_ensembles rejected the parameters")
    out.update(accum=accum, stats=stats)
    if want_buffers:
        out.update(counts=counts[:n], offsets=offsets[:n], depth_id=depth_id, total=total.value,
                   skipped=skipped.value)
    return out


def mirror_sorted(splats, sh_rest, sh_degree, cam, params, w, h):
    sp = np.ascontiguousarray(splats).view(np.float32).reshape(-1, 16)
    n = sp.shape[0]
    accum = np.zeros((h, w, 4), np.float32)
    stats = np.zeros(16, np.uint64)
    keys = np.zeros(n, np.uint32)
    vals = np.zeros(n, np.uint32)
    vis = u32(0)
    sh = np.ascontiguousarray(sh_rest if sh_rest is not None else np.zeros(1, np.float32), np.float32)
    rc = lib.mir_render_sorted(sp.ctypes.data_as(P(MGaussian)), n, ptr(sh, f32), sh_degree if sh_rest is not None else 0,
                               C.byref(camera_array(cam)), C.byref(params), w, h, params.spp_side, ptr(accum, f32),
                               ptr(stats, u64), ptr(keys, u32), ptr(vals, u32), C.byref(vis))
    if rc != 0:
        raise ValueError("mirror render_sorted failed")
    return dict(accum=accum, stats=stats, keys=keys, vals=vals, visible=vis.value)


def mirror_composite(accum, exposure, gamma_on, ow, oh):
    h, w = accum.shape[0], accum.shape[1]
    a = np.ascontiguousarray(accum, np.float32)
    out = np.zeros((oh, ow, 4), np.uint8)
    lib.mir_composite(ptr(a, f32), w, h, exposure, 1 if gamma_on else 0, ow, oh, ptr(out, C.c_uint8))
    return out


def mirror_fn(name, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    getattr(lib, name)(ptr(x, f32), ptr(y, f32), x.size)
    return y
