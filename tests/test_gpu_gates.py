"""The reference's GPU-vs-CPU verification gates (crates/gps-render/src/verify.rs), run on this renderer's GPU
images directly against the f64 gps-core oracle, with the reference's own thresholds: phase 2 (GPS on both
paths), phase 3 (the PBVR matrix, screen occupancy, view-conditioned), phase 4 (SH degrees), determinism and
history.  The bit-exact parity tests pin the GPU to the f32 mirror; these state the same result in verify.rs's
terms."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from scenes import BG, phase_scenes, pcg_py, to_gpu, to_image

import gps_hip as gh

pytestmark = pytest.mark.gpu
W = H = 96
CAM = gh.default_camera(200.0)
OCAM = O.default_camera(200.0)


def say(*a):
    print(*a, flush=True)


def params(**kw):
    pb = {k[5:]: kw.pop(k) for k in list(kw) if k.startswith("pbvr_")}
    return gh.RenderParams(spp_side=2, background=BG, pbvr=pb, **kw)


def gpu_mean(scene, p, sets, seed0=1, sh=None, deg=0):
    r = gh.GpsRenderer(W, H, p.spp_side)
    try:
        if sh is not None:
            r.set_gaussians_sh(scene, sh, deg)
        else:
            r.set_gaussians(scene)
        r.render_ensembles(CAM, p, seed0, sets)
        return to_image(r.read_accum()), r.read_stats()
    finally:
        r.close()


def within_4_sigma(a, b):
    return abs(float(a) - float(b)) <= 4.0 * max(math.sqrt(2.0 * float(b)), 1.0)


@pytest.mark.parametrize("idx", range(5))
def test_phase2_gps(idx):
    """verify.rs:92-144: both paths >= 60 dB against the Monte Carlo oracle, points within 4 sigma, >= 35 dB and
    SSIM >= 0.98 against the analytic image, no orphans or truncation, identical counters on both paths."""
    name, scene = phase_scenes()[idx]
    say(f"[phase2] {name}")
    ana = O.render_analytic(scene, OCAM, W, H, BG)
    seen = []
    for path in (0, 1):
        p = params(path=path)
        img, st = gpu_mean(to_gpu(scene), p, 300)
        opt = O.default_mc_options(per_particle_seed=path, max_points_per_splat=p.max_points_per_splat)
        cpu, cst = O.render_monte_carlo(scene, OCAM, W, H, 2, 300, BG, 1, opt)
        assert O.psnr(img, cpu) >= 60.0, name
        assert within_4_sigma(st.points, cst.points)
        assert st.orphan_subpixels == 0 and st.truncated_splats == 0
        assert O.psnr(img, ana) >= 35.0 and O.ssim(img, ana) >= 0.98
        seen.append(st)
    assert seen[0] == seen[1]


PBVR_CASES = [
    ("Proportional C0", 1, 0, 0, 0, 1), ("Proportional C3", 1, 3, 0, 0, 1), ("Extinction C0", 2, 0, 0, 0, 1),
    ("Extinction C1", 2, 1, 0, 0, 1), ("Extinction C2", 2, 2, 0, 0, 1), ("Extinction C3", 2, 3, 0, 0, 1),
    ("Extinction C3 no-jitter", 2, 3, 0, 0, 0), ("Extinction C3 centre-depth", 2, 3, 0, 1, 1),
    ("Extinction C3+R", 2, 3, 1, 0, 1), ("Extinction C3+R centre-depth", 2, 3, 1, 1, 1),
]


@pytest.mark.parametrize("si", [1, 3, 4])
@pytest.mark.parametrize("case", PBVR_CASES, ids=[c[0] for c in PBVR_CASES])
def test_phase3_pbvr(si, case):
    """verify.rs:368-403, 429-464: >= 50 dB against the f64 PBVR-3D oracle, points and candidates within
    4 sigma, C3+R >= 30 dB against the analytic image."""
    label, method, cal, radial, centre, jitter = case
    scene = phase_scenes()[si][1]
    say(f"[phase3] {label} scene {si}")
    p = params(method=method, pbvr_calibration=cal, pbvr_radial_correction=radial, pbvr_centre_depth=centre,
               pbvr_low_pass_jitter=jitter, max_points_per_splat=1 << 24)
    img, st = gpu_mean(to_gpu(scene), p, 120)
    opt = O.default_particle_options(level=cal, rule=0 if method == 1 else 1, radial_correction=radial,
                                     centre_depth=centre, low_pass_jitter=jitter, per_particle_seed=1)
    cpu, cst = O.render_particles_3d(scene, OCAM, W, H, 2, 120, BG, 1, opt)
    assert O.psnr(img, cpu) >= 50.0
    assert within_4_sigma(st.points, cst.points) and within_4_sigma(st.candidates, cst.candidates)
    assert st.orphan_subpixels == 0 and st.truncated_splats == 0
    if radial:
        assert O.psnr(img, O.render_analytic(scene, OCAM, W, H, BG)) >= 30.0


@pytest.mark.parametrize("si", [1, 3, 4])
def test_phase3_occupancy_and_view_conditioned(si):
    """verify.rs:467-516, 532-539"""
    scene = phase_scenes()[si][1]
    ana = O.render_analytic(scene, OCAM, W, H, BG)
    img, st = gpu_mean(to_gpu(scene), params(method=4, max_points_per_splat=1), 120)
    assert O.psnr(img, ana) >= 30.0 and st.orphan_subpixels == 0 and st.truncated_splats == 0
    assert st.candidates <= len(scene) * W * H * 4 * 120
    vc = params(method=3, pbvr_calibration=0, pbvr_base_k=4096.0, max_points_per_splat=1 << 24)
    img, st_vc = gpu_mean(to_gpu(scene), vc, 120)
    _, st_gps = gpu_mean(to_gpu(scene), params(max_points_per_splat=1 << 24), 120)
    assert O.psnr(img, ana) >= 30.0
    assert within_4_sigma(st_vc.points, st_gps.points) and st_vc.undercovered_splats == 0
    _, st_low = gpu_mean(to_gpu(phase_scenes()[1][1]), params(method=3, pbvr_calibration=0,
                                                               max_points_per_splat=1 << 24), 8)
    assert st_low.undercovered_splats > 0


def test_phase4_sh_degrees():
    """verify.rs:656-732: each SH degree >= 60 dB against the f64 eval_sh colours through the Monte Carlo oracle,
    and each degree changes the image (< 45 dB between consecutive degrees)."""
    stride = 15
    seed = 99

    def rnd():
        nonlocal seed
        seed = pcg_py(seed)
        return seed / 4294967296.0 - 0.5

    dcs, rests, base = [], [], []
    for i in range(4):
        dcs.append(np.array([rnd(), rnd(), rnd()]) * 1.2)
        rests.append(np.array([rnd() * 0.8 for _ in range(3 * stride)]))
        base.append(dict(pos=(-0.9 + 0.6 * i, 0.25 * (i - 1.5), 5.0 + 0.7 * i), log_scale=(math.log(0.22),) * 3,
                         opacity=0.85, color=(1, 1, 1), rot_xyzw=(0.0, 0.0, 0.0, 1.0)))
    sp = to_gpu(base)
    for i in range(4):
        c = dcs[i] * 0.28209479177387814 + 0.5
        sp[i]["color"] = (c[0], c[1], c[2], 0.0)
    rest32 = np.concatenate(rests).astype(np.float32)
    images = []
    for deg in range(4):
        expected = []
        for i, g in enumerate(base):
            d = np.array(g["pos"]) / np.linalg.norm(g["pos"])
            out = (O.f64 * 3)()
            r64 = np.ascontiguousarray(rests[i], np.float64)
            O.lib.orc_eval_sh(deg, (O.f64 * 3)(*dcs[i]), O.ptr(r64, O.f64), (O.f64 * 3)(*d), stride, out)
            expected.append(dict(g, color=tuple(out)))
        img, _ = gpu_mean(sp, params(sh_degree=deg), 300, sh=rest32, deg=3)
        cpu, _ = O.render_monte_carlo(expected, OCAM, W, H, 2, 300, BG, 1, O.default_mc_options(per_particle_seed=1))
        assert O.psnr(img, cpu) >= 60.0, deg
        images.append(img)
    for d in range(1, 4):
        assert O.psnr(images[d], images[d - 1]) < 45.0


def test_determinism_and_history():
    """verify.rs:560-605: the same seeds give the same image; a history built in two calls equals one call;
    different seeds differ."""
    scene = to_gpu(phase_scenes()[3][1])
    p = params(method=2, pbvr_radial_correction=1)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(scene)
    r.render_ensembles(CAM, p, 5, 7)
    a = r.read_accum()
    r.reset_accum()
    r.render_ensembles(CAM, p, 5, 3)
    r.render_ensembles(CAM, p, 8, 4)
    b = r.read_accum()
    r.reset_accum()
    r.render_ensembles(CAM, p, 50, 7)
    c = r.read_accum()
    r.close()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(a, c)


# ------------------------------------------------------------------ phase 2: composite, one huge splat, 33.5M subpixels
def _mirror_params(p):
    import ctypes
    m = O.MParams()
    ctypes.memmove(ctypes.byref(m), ctypes.byref(p), ctypes.sizeof(O.MParams))
    return m


def _ocam(c):
    import ctypes
    o = O.OrcCamera()
    ctypes.memmove(ctypes.byref(o), ctypes.byref(c), ctypes.sizeof(O.OrcCamera))
    return o


STAT_KEYS = ("points", "truncated_splats", "dropped_points", "visible_splats", "orphan_subpixels", "candidates",
             "undercovered_splats", "skipped_ensembles", "ensembles")


def _stats(st):
    return tuple(int(getattr(st, k)) for k in STAT_KEYS)


def test_phase2_composite_gates():
    """verify.rs:225-291 on the GPU composite: scene 1 after 16 ensembles; the five probes of the same-size composite
    and every pixel of the 2x upsampled one within 1.01 of the f64 single-OETF expectation (and within the 0.51 an
    RGBA8 target's rounding to nearest allows), also with exposure and gamma off and at odd sizes."""
    from composite_gate import REFERENCE_BOUND, STRICT, probe_error, worst_error
    scene = to_gpu(phase_scenes()[1][1])
    p = params()
    r = gh.GpsRenderer(W, H, 2)
    try:
        r.set_gaussians(scene)
        r.render_ensembles(CAM, p, 1, 16)
        acc = r.read_accum()
        assert (acc[..., 3] == 16.0).all()
        same = r.composite(p, W, H)
        say(f"[composite] probes {probe_error(same, acc):.3f}, all {worst_error(same, acc):.3f}")
        assert probe_error(same, acc) <= REFERENCE_BOUND
        assert worst_error(same, acc) <= STRICT
        up = r.composite(p, 2 * W, 2 * H)
        say(f"[upsample 2x] {worst_error(up, acc):.3f}")
        assert worst_error(up, acc) <= REFERENCE_BOUND and worst_error(up, acc) <= STRICT
        for ow, oh, gamma, expo in ((W, H, 0, 1.7), (50, 70, 1, 0.8), (3 * W, H + 1, 1, 1.2)):
            img = r.composite(p.copy(srgb_output=gamma, exposure=expo), ow, oh)
            assert worst_error(img, acc, expo, bool(gamma)) <= STRICT, (ow, oh, gamma, expo)
            assert (img[..., 3] == 255).all()
    finally:
        r.close()


def test_phase2_4m_particles_from_one_splat():
    """verify.rs:192-199: one splat (log-scale ln 0.55, o = 0.9) at density 400 and cap 2^24 gives more than 4M particles
    in one ensemble.  Particle path, 2 ensembles: >= 60 dB against the f64 Monte Carlo oracle with per-particle seeds,
    points within 4 sigma, no orphans; and bit-exact against the f32 mirror (image, counters, and the last
    ensemble's counts, offsets and depth|id buffer through the debug hook)."""
    s = dict(phase_scenes()[2][1][0], log_scale=(math.log(0.55),) * 3)
    scene = [s]
    p = params(path=1, density_scale=400.0, max_points_per_splat=1 << 24)
    gpu = {}
    for keep in (False, True):
        r = gh.GpsRenderer(W, H, 2)
        try:
            r.set_gaussians(to_gpu(scene))
            if keep:
                r.set_debug(gh.DEBUG_KEEP_BUFFERS)
            r.render_ensembles(CAM, p, 1, 2)
            gpu[keep] = dict(accum=r.read_accum(), stats=r.read_stats())
            if keep:
                gpu[keep].update(r.read_counts(), depth_id=r.read_depth_id())
        finally:
            r.close()
    st = gpu[False]["stats"]
    say(f"[4M+ one splat] points {st.points}, truncated {st.truncated_splats}")
    assert st.points > 4_000_000 and st.truncated_splats == 0 and st.orphan_subpixels == 0
    img = to_image(gpu[False]["accum"])
    opt = O.default_mc_options(per_particle_seed=1, density_scale=400.0, max_points_per_splat=1 << 24)
    cpu, cst = O.render_monte_carlo(scene, OCAM, W, H, 2, 2, BG, 1, opt)
    assert O.psnr(img, cpu) >= 60.0
    assert within_4_sigma(st.points, cst.points)
    m = O.mirror_render(to_gpu(scene), None, 0, OCAM, _mirror_params(p), W, H, 1, 2)
    last = O.mirror_render(to_gpu(scene), None, 0, OCAM, _mirror_params(p), W, H, 2, 1, want_buffers=True)
    for keep in (False, True):
        assert _stats(gpu[keep]["stats"]) == tuple(int(x) for x in m["stats"][:9])
        assert np.array_equal(gpu[keep]["accum"].view(np.uint32), m["accum"].view(np.uint32))
    b = gpu[True]
    assert np.array_equal(b["counts"], last["counts"]) and np.array_equal(b["offsets"], last["offsets"])
    assert b["total"] == last["total"] > 4_000_000
    assert np.array_equal(b["depth_id"], last["depth_id"])


def test_phase2_2048x1024_spp16():
    """verify.rs:294-322: 2048 x 1024 at spp side 4 (33.5M subpixels) with one splat and focal 900: the corners keep
    the background, the centre is covered, no orphan subpixels; bit-exact against the f32 mirror, including the
    depth|id buffer."""
    w, h, s = 2048, 1024, 4
    scene = to_gpu([phase_scenes()[1][1][0]])
    cam = gh.default_camera(900.0)
    p = gh.RenderParams(spp_side=s, background=BG)
    r = gh.GpsRenderer(96, 96, 2)
    try:
        r.resize(w, h, s)
        r.set_gaussians(scene)
        r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.reset_stats()
        r.render_ensembles(cam, p, 1, 1)
        acc = r.read_accum()
        st = r.read_stats()
        d = r.read_depth_id()
    finally:
        r.close()
    say(f"[2048x1024 spp16] points {st.points}")
    assert abs(acc[0, 0, 0] - BG[0]) < 1e-6 and abs(acc[-1, -1, 2] - BG[2]) < 1e-6
    assert abs(acc[h // 2, w // 2, 0] - BG[0]) > 1e-3
    assert st.orphan_subpixels == 0 and d.size == w * h * s * s
    m = O.mirror_render(scene, None, 0, _ocam(cam), _mirror_params(p), w, h, 1, 1, want_buffers=True)
    assert _stats(st) == tuple(int(x) for x in m["stats"][:9])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))
    assert np.array_equal(d, m["depth_id"])
