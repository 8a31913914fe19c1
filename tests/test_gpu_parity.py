"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

Bit-exact: device f32 functions, phase-0 packed atomics (crates/gps-render/src/lib.rs:99-192), the scan
(lib.rs:196-275), per-splat particle counts, exclusive offsets, particle totals, depth|id buffers,
accumulation buffers and every statistics counter of render_ensembles, against oracle/gps_mirror.c on the
same seeded inputs.
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from scenes import BG, crowd, many_splats, phase_scenes, to_gpu

pytestmark = pytest.mark.gpu

gh = pytest.importorskip("gps_hip")
from gps_hip import scenes as S  # noqa: E402

STAT_KEYS = ("points", "truncated_splats", "dropped_points", "visible_splats", "orphan_subpixels", "candidates",
             "undercovered_splats", "skipped_ensembles", "ensembles")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mparams(p):
    """gps_hip.RenderParams -> oracle MParams (same field order)."""
    m = O.MParams()
    import ctypes
    ctypes.memmove(ctypes.byref(m), ctypes.byref(p), ctypes.sizeof(O.MParams))
    return m


def ocam(c):
    o = O.OrcCamera()
    import ctypes
    ctypes.memmove(ctypes.byref(o), ctypes.byref(c), ctypes.sizeof(O.OrcCamera))
    return o


def stats_tuple(st):
    return tuple(int(getattr(st, k)) for k in STAT_KEYS)


def mirror_stats_tuple(m):
    return tuple(int(m[i]) for i in range(9))


# ------------------------------------------------------------------ elementary functions
@pytest.mark.parametrize("fn,lo,hi", [("exp", -104.0, 89.0), ("log", -30.0, 30.0), ("sin", 0.0, 6.2832),
                                      ("cos", 0.0, 6.2832)])
def test_device_math_bit_exact(fn, lo, hi):
    say(f"[math] {fn}")
    rng = np.random.default_rng(3)
    if fn == "log":
        x = np.exp(rng.uniform(lo, hi, 1_000_000)).astype(np.float32)
        x = np.concatenate([x, rng.uniform(0.0, 2.0, 1_000_000).astype(np.float32),
                            np.array([0.0, 1e-45, 1.0, np.inf], np.float32)])
    else:
        x = rng.uniform(lo, hi, 2_000_000).astype(np.float32)
    dev = gh.math_eval(fn, x)
    ref = O.mirror_fn(f"mir_{fn}f_n", x)
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))


def test_device_sqrt_div_correctly_rounded():
    say("[math] sqrt/div")
    rng = np.random.default_rng(4)
    x = np.exp(rng.uniform(-80, 80, 1_000_000)).astype(np.float32)
    y = np.exp(rng.uniform(-80, 80, 1_000_000)).astype(np.float32)
    assert np.array_equal(gh.math_eval("sqrt", x).view(np.uint32), np.sqrt(x).view(np.uint32))
    assert np.array_equal(gh.math_eval("div", x, y).view(np.uint32), (x / y).view(np.uint32))


# ------------------------------------------------------------------ phase 0
def test_phase0_packed_atomic_equals_two_pass():
    """p0_sample.wgsl two-pass depth/colour atomics == one packed 64-bit minimum (lib.rs:169-191)."""
    PIX, SAMPLES = 4096, 200_000
    for e in range(3):
        seed = 1000 + e * 7
        say(f"[phase0 atomic] seed={seed}")
        d, c = gh.phase0_atomic(seed, PIX, SAMPLES)
        rd = np.zeros(PIX, np.uint32)
        rc = np.zeros(PIX, np.uint32)
        O.lib.mir_phase0(seed, 0, PIX, SAMPLES, O.ptr(rd, O.u32), O.ptr(rc, O.u32))
        assert np.array_equal(d, rd) and np.array_equal(c, rc)


@pytest.mark.parametrize("n", [1, 255, 256, 257, 65_536, 300_000, 1_000_003, 40_000_000])
def test_scan_exact(n):
    """lib.rs:196-275 with input pcg_hash(i) % 7."""
    say(f"[scan] n={n}")
    i = np.arange(n, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    s = (i * np.uint64(747796405) + np.uint64(2891336453)) & m
    w = (((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737)) & m
    x = (((w >> np.uint64(22)) ^ w) % np.uint64(7)).astype(np.uint32)
    got = gh.scan_exclusive_u32(x)
    ref = np.concatenate([[0], np.cumsum(x[:-1], dtype=np.uint64)]).astype(np.uint32)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [2049, 3_000_017])
def test_scan_wraps_modulo_2_32(n):
    """Full-range u32 inputs: prefix sums pass 2^36 within a few tiles, and the output is the exclusive sum modulo
    2^32 (u32 arithmetic, gps_scan.wgsl:11-36)."""
    say(f"[scan wrap] n={n}")
    x = np.random.default_rng(n).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    got = gh.scan_exclusive_u32(x)
    ref = np.concatenate([np.zeros(1, np.uint64), np.cumsum(x[:-1], dtype=np.uint64)]) & np.uint64(0xFFFFFFFF)
    assert int(np.sum(x, dtype=np.uint64)) > (1 << 36) or n < 100_000
    assert np.array_equal(got, ref.astype(np.uint32))


# ------------------------------------------------------------------ render_ensembles vs the mirror
W = H = 96


def run_pair(scene_arr, sh, deg, cam, p, w, h, first_seed, count, keep=True):
    """Render `count` ensembles on the GPU and in the mirror; return both results."""
    r = gh.GpsRenderer(w, h, p.spp_side)
    try:
        if sh is not None:
            r.set_gaussians_sh(scene_arr, sh, deg)
        else:
            r.set_gaussians(scene_arr)
        if keep:
            r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.reset_stats()
        r.render_ensembles(cam, p, first_seed, count)
        acc = r.read_accum()
        st = r.read_stats()
        bufs = None
        if keep and (p.path == 1 or p.method != 0):
            bufs = r.read_counts()
            bufs["depth_id"] = r.read_depth_id()
    finally:
        r.close()
    m = O.mirror_render(scene_arr, sh, deg, ocam(cam), mparams(p), w, h, first_seed, count, want_buffers=keep)
    return dict(accum=acc, stats=st, bufs=bufs), m


def assert_exact(gpu, m, check_buffers=True):
    assert stats_tuple(gpu["stats"]) == mirror_stats_tuple(m["stats"])
    assert np.array_equal(gpu["accum"].view(np.uint32), m["accum"].view(np.uint32))
    if check_buffers and gpu["bufs"] is not None:
        b = gpu["bufs"]
        assert b["skipped"] == m["skipped"] and b["total"] == m["total"]
        assert np.array_equal(b["offsets"], m["offsets"])
        if not m["skipped"]:
            assert np.array_equal(b["counts"], m["counts"])
        assert np.array_equal(b["depth_id"], m["depth_id"])


CAM = gh.default_camera(200.0)


@pytest.mark.parametrize("idx", range(5))
@pytest.mark.parametrize("path", [0, 1])
def test_gps_bit_exact(idx, path):
    name, scene = phase_scenes()[idx]
    say(f"[gps] {name} path={path}")
    p = gh.RenderParams(spp_side=2, background=BG, path=path)
    gpu, m = run_pair(to_gpu(scene), None, 0, CAM, p, W, H, 1, 6)
    assert_exact(gpu, m)
    assert gpu["stats"].orphan_subpixels == 0


PBVR_CASES = [(1, 0, 0, 0, 1), (1, 3, 0, 0, 1), (2, 0, 0, 0, 1), (2, 1, 0, 0, 1), (2, 2, 0, 0, 1), (2, 3, 0, 0, 1),
              (2, 3, 0, 0, 0), (2, 3, 0, 1, 1), (2, 3, 1, 0, 1), (2, 3, 1, 1, 1), (3, 0, 0, 0, 1)]


@pytest.mark.parametrize("case", PBVR_CASES)
@pytest.mark.parametrize("si", [1, 3, 4])
def test_pbvr_bit_exact(case, si):
    method, cal, radial, centre, jitter = case
    name, scene = phase_scenes()[si]
    say(f"[pbvr] {name} case={case}")
    p = gh.RenderParams(spp_side=2, background=BG, method=method, max_points_per_splat=1 << 24,
                        pbvr=dict(calibration=cal, radial_correction=radial, centre_depth=centre,
                                  low_pass_jitter=jitter))
    gpu, m = run_pair(to_gpu(scene), None, 0, CAM, p, W, H, 1, 4)
    assert_exact(gpu, m)


@pytest.mark.parametrize("density", [0.25, 2.0])
@pytest.mark.parametrize("si", [1, 3, 4])
def test_occupancy_bit_exact(density, si):
    name, scene = phase_scenes()[si]
    say(f"[occupancy] {name} density={density}")
    p = gh.RenderParams(spp_side=2, background=BG, method=4, density_scale=density, max_points_per_splat=1)
    gpu, m = run_pair(to_gpu(scene), None, 0, CAM, p, W, H, 500, 4)
    assert_exact(gpu, m)


def test_cap_truncation_and_exact_radius():
    scene = phase_scenes()[2][1]
    for kw in (dict(max_points_per_splat=50), dict(exact_radius=1), dict(spp_side=1), dict(spp_side=3)):
        say(f"[gps variant] {kw}")
        p = gh.RenderParams(spp_side=2, background=BG).copy(**kw)
        gpu, m = run_pair(to_gpu(scene), None, 0, CAM, p, W, H, 7, 3)
        assert_exact(gpu, m)


def test_sh_degrees_bit_exact():
    rng = np.random.default_rng(5)
    scene = S.small_uniform(512, seed=3)
    sh = rng.normal(0, 0.3, 512 * 3 * 15).astype(np.float32)
    cam = S.small_uniform_camera(64, 64)
    for deg in range(4):
        say(f"[sh] degree {deg}")
        p = gh.RenderParams(spp_side=1, sh_degree=deg, method=2)
        gpu, m = run_pair(scene, sh, 3, cam, p, 64, 64, 3, 3)
        assert_exact(gpu, m)


def test_many_splats_three_levels():
    """verify.rs:173-186 and 541-557: 70,000 splats (several scan tiles)."""
    scene = to_gpu(many_splats())
    for p in (gh.RenderParams(spp_side=2, background=BG),
              gh.RenderParams(spp_side=2, background=BG, method=2, max_points_per_splat=1 << 24,
                              pbvr=dict(radial_correction=1))):
        say(f"[70k] method={p.method}")
        gpu, m = run_pair(scene, None, 0, CAM, p, W, H, 1, 3)
        assert_exact(gpu, m)


def test_small_uniform_extinction_c3_16_ensembles():
    """4,096-Gaussian 256x256 PBVR Extinction C3 workload: 16 ensembles, seeds 1..16."""
    say("[small uniform] 16 ensembles")
    scene = S.small_uniform()
    cam = S.small_uniform_camera()
    p = gh.RenderParams(spp_side=1, method=2, max_points_per_splat=16384)
    gpu, m = run_pair(scene, None, 0, cam, p, 256, 256, 1, 16)
    assert_exact(gpu, m)
    # batched (no debug hook) must equal the per-ensemble run
    r = gh.GpsRenderer(256, 256, 1)
    r.set_gaussians(scene)
    r.render_ensembles(cam, p, 1, 16)
    acc = r.read_accum()
    r.close()
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))


def test_overflow_skip_exact():
    scene = phase_scenes()[3][1]
    p = gh.RenderParams(spp_side=2, background=BG, density_scale=1.0e7, max_points_per_splat=0xFFFFFFFF)
    say("[overflow]")
    gpu, m = run_pair(to_gpu(scene), None, 0, CAM, p, 16, 16, 1, 1)
    assert m["skipped"] == 1 and gpu["bufs"]["skipped"] == 1
    assert_exact(gpu, m)


def test_offscreen_camera_is_background():
    """verify.rs:587-591"""
    scene = to_gpu(phase_scenes()[3][1])
    cam = gh.default_camera(200.0)
    cam.view_pos[:] = [1000.0, 0.0, 0.0]
    p = gh.RenderParams(spp_side=2, background=BG, method=2, pbvr=dict(radial_correction=1))
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(scene)
    r.render_ensembles(cam, p, 800, 2)
    acc = r.read_accum()
    r.close()
    assert (acc[..., 3] == 2.0).all()
    for k in range(3):
        assert np.abs(acc[..., k] - 2.0 * np.float32(BG[k])).max() < 1e-6


def test_history_and_batched_vs_separate():
    """verify.rs:212-223, 572-605: counts, reset, batched == separate, determinism."""
    scene = to_gpu(phase_scenes()[3][1])
    p = gh.RenderParams(spp_side=2, background=BG, method=2, max_points_per_splat=1 << 24,
                        pbvr=dict(radial_correction=1))
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(scene)
    r.render_ensembles(CAM, p, 700, 8)
    expected = r.read_accum()
    assert (expected[..., 3] == 8.0).all() and r.accumulated() == 8
    r.reset_accum()
    assert (r.read_accum() == 0).all() and r.accumulated() == 0
    for s in range(700, 708):
        r.render_ensembles(CAM, p, s, 1)
    assert np.array_equal(r.read_accum(), expected)
    r.reset_accum()
    r.render_ensembles(CAM, p, 700, 4)
    r.render_ensembles(CAM, p, 704, 4)
    assert np.array_equal(r.read_accum(), expected)
    r.reset_accum()
    r.render_ensembles(CAM, p, 600, 8)
    assert not np.array_equal(r.read_accum(), expected)
    r.close()


def test_composite_bit_exact():
    scene = to_gpu(phase_scenes()[1][1])
    p = gh.RenderParams(spp_side=2, background=BG)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(scene)
    r.render_ensembles(CAM, p, 1, 16)
    acc = r.read_accum()
    for ow, oh, gamma, expo in ((W, H, 1, 1.0), (2 * W, 2 * H, 1, 1.0), (W, H, 0, 1.7), (50, 70, 1, 0.8)):
        q = p.copy(srgb_output=gamma, exposure=expo)
        img = r.composite(q, ow, oh)
        ref = O.mirror_composite(acc, expo, gamma, ow, oh)
        assert np.array_equal(img, ref), (ow, oh, gamma, expo)
    r.close()


def test_empty_scene_renders_background():
    r = gh.GpsRenderer(32, 32, 2)
    p = gh.RenderParams(spp_side=2, background=BG)
    r.render_ensembles(CAM, p, 1, 3)
    acc = r.read_accum()
    r.close()
    assert (acc[..., 3] == 3.0).all()
    assert np.abs(acc[..., 0] - 3 * np.float32(BG[0])).max() < 1e-6


def test_errors_are_reported_not_fatal():
    r = gh.GpsRenderer(32, 32, 2)
    with pytest.raises(gh.GpsError):
        r.render_ensembles(CAM, gh.RenderParams(spp_side=1), 1, 1)
    with pytest.raises(gh.GpsError):
        r.set_gaussians_sh(to_gpu(phase_scenes()[1][1]), np.zeros(5, np.float32), 3)
    r.render_ensembles(CAM, gh.RenderParams(spp_side=2), 1, 1)
    r.close()
    with pytest.raises(gh.GpsError):
        gh.GpsRenderer(0, 10, 1)


# ------------------------------------------------------------------ sorted path (verify.rs:751-890)
def sorted_scenes():
    out = [(n, s) for n, s in phase_scenes()]
    out.append(("crowd of 3000", crowd()))
    return out


@pytest.mark.parametrize("idx", range(6))
def test_sorted_path_vs_mirror_and_analytic(idx):
    name, scene = sorted_scenes()[idx]
    say(f"[sorted] {name}")
    arr = to_gpu(scene)
    p = gh.RenderParams(spp_side=2, background=BG)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(arr)
    r.render_sorted(CAM, p, 1)
    acc = r.read_accum()
    keys, vals = r.read_sorted()
    st = r.read_stats()
    assert r.accumulated() == 1
    r.close()
    m = O.mirror_sorted(arr, None, 0, ocam(CAM), mparams(p), W, H)
    assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
    assert st.probe_demand == int(m["stats"][9])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))
    img = acc[..., :3].astype(np.float64)
    ana = O.render_analytic(scene, ocam(CAM), W, H, BG)
    assert O.psnr(img, ana) >= 40.0 and O.ssim(img, ana) >= 0.99 and (acc[..., 3] == 1.0).all()
    assert (np.diff(keys.astype(np.int64)) >= 0).all()
    assert np.array_equal(np.sort(vals), np.arange(len(vals)))


@pytest.mark.parametrize("seed", [1, 2])
def test_sorted_blend_reach_skip_with_needles(seed):
    """Needles, whose minor axis sits at the 1e-12 eigenvalue clamp, and discs at random orientations and depths,
    sub-pixel to frame-filling, opacities near 1/255 and near 1: keys, values and the accumulation buffer of the
    sorted alpha path (tile lists of the binned quads) equal the oracle's, which blends every splat at every pixel."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(700):
        big = rng.uniform(-4.0, 0.5)
        thin = big - rng.uniform(0.0, 14.0) if i % 3 else big - rng.uniform(0.0, 1.0)
        ls = [big, thin, rng.uniform(thin, big)]
        rng.shuffle(ls)
        q = rng.normal(size=4)
        out.append(dict(pos=(rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(3.0, 9.0)),
                        log_scale=tuple(ls), rot_xyzw=tuple(q / np.linalg.norm(q)),
                        opacity=float(rng.choice([0.0045, 0.02, 0.5, 0.98])), color=tuple(rng.uniform(0, 1, 3))))
    arr = to_gpu(out)
    say(f"[sorted needles] seed {seed}")
    p = gh.RenderParams(spp_side=1, background=BG)
    r = gh.GpsRenderer(W, H, 1)
    r.set_gaussians(arr)
    r.render_sorted(CAM, p, 1)
    acc = r.read_accum()
    keys, vals = r.read_sorted()
    r.close()
    m = O.mirror_sorted(arr, None, 0, ocam(CAM), mparams(p), W, H)
    assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))
    assert (acc[..., :3] != np.asarray(BG[:3], np.float32)).any(axis=-1).mean() > 0.3


@pytest.mark.parametrize("view", ["near", "mid", "far"])
def test_sorted_clustered_sh3_views_match_oracle(view):
    """The sorted alpha path on a clustered SH3 scene (12,000 Gaussians, seed 5: clusters of small splats among
    large ones, every size and orientation) at 320 x 180 from the three orbit views: tile lists of hundreds of
    splats over many 64-splat batches, most splats culled per 8x8 quadrant by the binning box and the axis test.
    Keys, values and the accumulation buffer equal the oracle's, which blends every splat at every pixel."""
    w, h = 320, 180
    g, sh = S.clustered(12_000, seed=5, sh_degree=3)
    cam = S.clustered_camera(view, w, h)
    say(f"[sorted clustered] {view}")
    p = gh.RenderParams(spp_side=1, background=BG)
    r = gh.GpsRenderer(w, h, 1)
    r.set_gaussians_sh(g, sh, 3)
    r.render_sorted(cam, p, 1)
    acc = r.read_accum()
    keys, vals = r.read_sorted()
    r.close()
    m = O.mirror_sorted(g, sh, 3, ocam(cam), mparams(p), w, h)
    assert m["visible"] > 1000
    assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))


@pytest.mark.parametrize("seed", [21, 22, 23, 24])
def test_sorted_random_views_match_oracle(seed):
    """Random orbit views (distance 1 to 12, any yaw, pitch within +-60 degrees) of a 6,000-splat clustered SH3 scene
    at 256 x 144, some from inside the clusters: keys, values and the image of the sorted path equal the oracle's."""
    rng = np.random.default_rng(seed)
    w, h = 256, 144
    g, sh = S.clustered(6_000, seed=seed, sh_degree=3)
    cam = gh.orbit_camera((0.0, 0.0, 0.0), float(rng.uniform(1.0, 12.0)), float(rng.uniform(0.0, 2.0 * math.pi)),
                          float(rng.uniform(-1.05, 1.05)), (0.0, 0.0, 1.0), S.FOV_60, w, h)
    p = gh.RenderParams(spp_side=1, background=BG)
    r = gh.GpsRenderer(w, h, 1)
    r.set_gaussians_sh(g, sh, 3)
    r.render_sorted(cam, p, 1)
    acc = r.read_accum()
    keys, vals = r.read_sorted()
    r.close()
    m = O.mirror_sorted(g, sh, 3, ocam(cam), mparams(p), w, h)
    say(f"[sorted random view] seed {seed}: {m['visible']} visible")
    assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))


def test_sorted_empty_frames_and_buffer_reuse():
    """One renderer through frames of the sorted path on scenes of different sizes, some seen from behind (no splat
    visible, no tile pairs): every frame's keys, values and accumulation buffer equal the oracle's, whatever the
    buffers the previous frames sized, and a frame with nothing visible is the background."""
    p = gh.RenderParams(spp_side=1, background=BG)
    back = gh.default_camera(200.0)
    back.view_rot[:] = [-1, 0, 0, 0, 1, 0, 0, 0, -1]   # turned half a turn about y: every splat behind
    big, small, other = crowd(), crowd(1200, seed=7), phase_scenes()[1][1]
    r = gh.GpsRenderer(W, H, 1)
    for i, (scene, cam) in enumerate([(big, CAM), (big, back), (other, CAM), (small, back), (small, CAM),
                                      (big, CAM)]):
        say(f"[sorted reuse] frame {i}")
        arr = to_gpu(scene)
        r.set_gaussians(arr)
        r.render_sorted(cam, p, 1)
        acc = r.read_accum()
        keys, vals = r.read_sorted()
        m = O.mirror_sorted(arr, None, 0, ocam(cam), mparams(p), W, H)
        assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
        assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))
        if cam is back:
            assert m["visible"] == 0
            assert (acc[..., :3] == np.asarray(BG, np.float32)).all()
        else:
            assert m["visible"] > 0
    r.close()


def test_sorted_large_frame_matches_oracle():
    """A 2000 x 1100 frame of the sorted path: 125 x 69 tiles (the last row partial), tile ids of 14 bits sorted in
    two 7-bit passes, 34,500 quadrant workgroups; keys, values and the image equal the oracle's."""
    w, h = 2000, 1100
    arr = to_gpu(crowd(600, seed=11))
    cam = gh.default_camera(1100.0)
    p = gh.RenderParams(spp_side=1, background=BG)
    r = gh.GpsRenderer(w, h, 1)
    r.set_gaussians(arr)
    r.render_sorted(cam, p, 1)
    acc = r.read_accum()
    keys, vals = r.read_sorted()
    r.close()
    m = O.mirror_sorted(arr, None, 0, ocam(cam), mparams(p), w, h)
    assert m["visible"] == 600
    assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))
    assert (acc[..., :3] != np.asarray(BG, np.float32)).any(axis=-1).mean() > 0.2


def test_radix_sort_300k_multilevel():
    """verify.rs:827-853: 300k small splats."""
    say("[sorted] 300k")
    rng = np.random.default_rng(11)
    n = 300_000
    scene = np.zeros(n, gh.GAUSSIAN_DTYPE)
    scene["pos_opacity"][:, 0] = rng.uniform(-1.3, 1.3, n)
    scene["pos_opacity"][:, 1] = rng.uniform(-1.3, 1.3, n)
    scene["pos_opacity"][:, 2] = rng.uniform(4.0, 9.0, n)
    scene["pos_opacity"][:, 3] = 0.5
    scene["log_scale"][:, :3] = math.log(0.01)
    scene["rot_wxyz"][:, 0] = 1.0
    scene["color"][:, :3] = rng.uniform(0, 1, (n, 3))
    p = gh.RenderParams(spp_side=2, background=BG)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(scene)
    r.render_sorted(CAM, p, 1)
    keys, vals = r.read_sorted()
    acc = r.read_accum()
    r.close()
    m = O.mirror_sorted(scene, None, 0, ocam(CAM), mparams(p), W, H)
    assert np.array_equal(keys, m["keys"]) and np.array_equal(vals, m["vals"])
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))


def test_demand_probe_matches_gps_points():
    """verify.rs:855-877: probe within 2 % of the GPS path's mean points per ensemble."""
    arr = to_gpu(crowd())
    p = gh.RenderParams(spp_side=2, background=BG, max_points_per_splat=1 << 24)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(arr)
    r.reset_stats()
    r.probe_demand(CAM, p)
    probe = r.read_stats().probe_demand
    r.reset_stats()
    r.render_ensembles(CAM, p, 1, 40)
    pts = r.read_stats().points / 40.0
    r.close()
    assert pts > 1000.0 and abs(probe - pts) / pts < 0.02


def test_sorted_after_resize():
    arr = to_gpu(crowd())
    p = gh.RenderParams(spp_side=2, background=BG)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians(arr)
    r.render_sorted(CAM, p, 1)
    r.resize(64, 64, 2)
    r.set_gaussians(arr)
    r.render_sorted(CAM, p, 1)
    assert r.read_accum().shape == (64, 64, 4)
    r.close()


# ------------------------------------------------------------------ 64-bit onesweep sort and the depth-sorted resolve
@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 100_003, 1_000_000, 20_000_000])
def test_onesweep_sort_stable(n):
    """Stable sort of (u64 key, u32 value) = numpy's stable argsort, with duplicate keys and constant bytes."""
    say(f"[onesweep] n={n}")
    rng = np.random.default_rng(n)
    hi = (np.uint64(0xBF) << np.uint64(56)) | (rng.integers(0, 1 << 12, n, dtype=np.uint64) << np.uint64(40))
    keys = hi | rng.integers(0, max(n // 3, 1), n, dtype=np.uint64)     # many duplicates
    vals = np.arange(n, dtype=np.uint32)
    k, v = gh.sort_pairs_u64(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


@pytest.mark.parametrize("items", [8, 32, 64])
@pytest.mark.parametrize("n", [8191, 8193, 16385, 1_000_003])
def test_onesweep_sort_tiles(items, n):
    """Every sort tile size (256 x 8, 256 x 16 default, 512 x 16, 1024 x 16) gives numpy's stable order."""
    rng = np.random.default_rng(n + items)
    keys = rng.integers(0, 1 << 20, n, dtype=np.uint64) << np.uint64(40) | rng.integers(0, 1 << 12, n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    saved = gh.get_tuning()
    try:
        gh.set_tuning(sort_items_per_thread=items)
        k2, v2 = gh.sort_pairs_u64(keys, vals)
    finally:
        gh.set_tuning(**saved)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k2, keys[order]) and np.array_equal(v2, vals[order])


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("n", [4097, 1_000_003])
def test_onesweep_sort_digit_widths(bits, n):
    """Every radix digit width (4..8 bits; 64 is not a multiple of 5, 6 or 7, so the last digit is narrower)
    gives numpy's stable order on full-range keys with duplicates and on keys with constant bytes."""
    rng = np.random.default_rng(n + bits)
    full = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    dup = full[3::7]
    full[: 7 * dup.size : 7] = dup   # duplicate keys
    narrow = (np.uint64(0xBF) << np.uint64(56)) | rng.integers(0, max(n // 5, 1), n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    saved = gh.get_tuning()
    try:
        gh.set_tuning(sort_digit_bits=bits)
        for keys in (full, narrow):
            k2, v2 = gh.sort_pairs_u64(keys, vals)
            order = np.argsort(keys, kind="stable")
            assert np.array_equal(k2, keys[order]) and np.array_equal(v2, vals[order])
    finally:
        gh.set_tuning(**saved)
    with pytest.raises(gh.GpsError):
        gh.set_tuning(sort_digit_bits=9)
    with pytest.raises(gh.GpsError):
        gh.set_tuning(sort_digit_bits=3)


@pytest.mark.parametrize("bits", [4, 6, 7])
def test_sorted_particle_path_digit_widths(bits):
    """The depth-sorted resolve equals the packed minimum with every sort digit width."""
    name, scene = phase_scenes()[3]
    p = gh.RenderParams(spp_side=2, background=BG, method=2, max_points_per_splat=1 << 24,
                        pbvr=dict(radial_correction=1))
    saved = gh.get_tuning()
    try:
        gh.set_tuning(sort_digit_bits=bits)
        sorted_vs_atomic(to_gpu(scene), None, 0, CAM, p, W, H, (3, 3))
    finally:
        gh.set_tuning(**saved)


def test_onesweep_sort_full_range_keys():
    n = 3_000_000
    rng = np.random.default_rng(9)
    keys = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    k, v = gh.sort_pairs_u64(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


def sorted_vs_atomic(scene_arr, sh, deg, cam, p, w, h, seeds):
    out = []
    for path in (1, 2):
        q = p.copy(path=path)
        r = gh.GpsRenderer(w, h, q.spp_side)
        if sh is not None:
            r.set_gaussians_sh(scene_arr, sh, deg)
        else:
            r.set_gaussians(scene_arr)
        r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.render_ensembles(cam, q, seeds[0], seeds[1])
        res = dict(accum=r.read_accum(), depth=r.read_depth_id(), stats=r.read_stats())
        if path == 2:
            res["keys"], res["sub"] = r.read_particle_keys()
            res["total"] = r.read_counts()["total"]
        r.close()
        out.append(res)
    a, b = out
    assert np.array_equal(a["depth"], b["depth"])
    assert np.array_equal(a["accum"].view(np.uint32), b["accum"].view(np.uint32))
    assert a["stats"] == b["stats"]
    k = b["keys"]
    assert k.size == b["total"]
    assert (np.diff(k.astype(np.uint64)) > 0).all() if k.size > 1 else True   # unique keys, ascending
    idx = (k & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.array_equal(np.sort(idx), np.arange(k.size))                   # particle indices: a permutation
    return b


@pytest.mark.parametrize("si", [1, 3, 4])
@pytest.mark.parametrize("method", [0, 2])
def test_sorted_particle_path_equals_atomic(si, method):
    name, scene = phase_scenes()[si]
    say(f"[sorted particles] {name} method={method}")
    p = gh.RenderParams(spp_side=2, background=BG, method=method, max_points_per_splat=1 << 24,
                        pbvr=dict(radial_correction=1))
    sorted_vs_atomic(to_gpu(scene), None, 0, CAM, p, W, H, (3, 3))


def test_sorted_particle_path_3m_splats_1080p():
    """3,000,000 clustered Gaussians (seed 2), 1920x1080, SPP side 2, GPS, cap 65,536, far camera: the
    depth-sorted resolve equals the packed-minimum path bit for bit, and the sorted keys are the unique
    ascending order of the generated particles."""
    say("[sorted particles] 3M")
    g, sh = S.clustered(3_000_000, seed=2, sh_degree=3)
    cam = S.clustered_camera("far", 1920, 1080)
    p = gh.RenderParams(spp_side=2, method=0, max_points_per_splat=65536)
    b = sorted_vs_atomic(g, sh, 3, cam, p, 1920, 1080, (1, 1))
    say(f"[sorted particles] 3M total particles {b['total']}")
    assert b["total"] > 10_000_000


@pytest.mark.timeout(900)
def test_config3_near_sorted_path_equals_packed_minimum_and_oracle():
    """Config 3 at its own camera: 3,000,000 clustered Gaussians (seed 2, SH3), 1920x1080, SPP side 2, GPS, cap
    65,536, near orbit camera; about 827M particles, so 827M 64-bit keys per ensemble.  The sorted keys are
    unique and ascending, their particle indices are a permutation of 0..n-1, and the depth-sorted resolve equals
    the packed-minimum path and the oracle bit for bit (depth|id buffer, image, counters)."""
    g, sh = S.clustered(3_000_000, seed=2, sh_degree=3)
    cam = S.clustered_camera("near", 1920, 1080)
    p = gh.RenderParams(spp_side=2, method=0, max_points_per_splat=65536)
    res = {}
    for path in (2, 1):
        say(f"[config3 near] path {path}")
        r = gh.GpsRenderer(1920, 1080, 2)
        try:
            r.set_gaussians_sh(g, sh, 3)
            r.set_debug(gh.DEBUG_KEEP_BUFFERS)
            r.render_ensembles(cam, p.copy(path=path), 1, 1)
            res[path] = dict(accum=r.read_accum(), depth=r.read_depth_id(), stats=r.read_stats(),
                             total=r.read_counts()["total"])
            if path == 2:
                keys, _ = r.read_particle_keys()
        finally:
            r.close()
    n = res[2]["total"]
    say(f"[config3 near] {n} sorted keys")
    assert n > 800_000_000 and keys.size == n
    step = 1 << 26
    for i in range(0, n - 1, step):                       # strictly ascending: unique keys
        j = min(i + step, n - 1)
        assert (keys[i + 1:j + 1] > keys[i:j]).all(), i
    seen = np.zeros(n, np.bool_)                          # n indices below n, all present: a permutation
    for i in range(0, n, step):
        idx = (keys[i:i + step] & np.uint64(0xFFFFFFFF)).astype(np.int64)
        assert idx.max() < n
        seen[idx] = True
    assert seen.all()
    del keys, seen
    a, b = res[1], res[2]
    assert np.array_equal(a["depth"], b["depth"])
    assert np.array_equal(a["accum"].view(np.uint32), b["accum"].view(np.uint32))
    assert a["stats"] == b["stats"]
    m = O.mirror_render(g, sh, 3, ocam(cam), mparams(p.copy(path=1)), 1920, 1080, 1, 1, want_buffers=True)
    assert m["total"] == n
    assert stats_tuple(a["stats"]) == mirror_stats_tuple(m["stats"])
    assert np.array_equal(a["accum"].view(np.uint32), m["accum"].view(np.uint32))
    assert np.array_equal(a["depth"], m["depth_id"])


# ------------------------------------------------------------------ launch tuning never changes results
def test_tunings_are_bit_identical():
    scene = S.small_uniform(4096, seed=0)
    cam = S.small_uniform_camera(128, 128)
    p = gh.RenderParams(spp_side=2, method=2, max_points_per_splat=16384, pbvr=dict(radial_correction=1))
    saved = gh.get_tuning()
    ref = None
    try:
        for waves, ppt, items, batch, block in ((8, 4, 16, 0, 256), (4, 1, 8, 3, 128), (16, 8, 32, 1, 512),
                                                (1, 2, 8, 7, 256), (2, 8, 64, 0, 128), (3, 4, 16, 2, 512)):
            say(f"[tuning] waves={waves} ppt={ppt} items={items} batch={batch} block={block}")
            gh.set_tuning(particle_waves_per_cu=waves, particles_per_thread=ppt, sort_items_per_thread=items,
                          batch_ensembles=batch, particle_block=block)
            for path in (1, 2):
                r = gh.GpsRenderer(128, 128, 2)
                r.set_gaussians(scene)
                r.render_ensembles(cam, p.copy(path=path), 5, 9)
                acc = r.read_accum()
                r.close()
                if ref is None:
                    ref = acc
                assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32))
    finally:
        gh.set_tuning(**saved)


def sparse_visible_scene(seed=3):
    """Mostly zero-particle splats (opacity 0: no particles) with visible ones at gaps of 1 to 12,000 positions,
    so the offsets hold long runs of equal values for the chunk-owner searches to cross."""
    rng = np.random.default_rng(seed)
    gaps = rng.choice([1, 3, 40, 70, 700, 5000, 12000], size=160)
    idx = np.cumsum(gaps)
    vis = S.small_uniform(len(idx), seed=seed)
    g = np.zeros(int(idx[-1]) + 50, vis.dtype)
    g["pos_opacity"][:, :3] = rng.uniform(-1.0, 1.0, (len(g), 3))
    g["log_scale"][:, :3] = -3.5
    g["rot_wxyz"][:, 0] = 1.0
    g["pos_opacity"][:, 3] = 0.0
    g[idx] = vis
    return g


@pytest.mark.parametrize("method", [0, 2])
def test_owner_search_across_zero_particle_runs(method):
    """Chunk owners found across long runs of zero-particle splats, with tiny chunks (one particle per thread,
    128-thread blocks) and runs of chunks that cross ensembles, bit-exact against the oracle."""
    scene = sparse_visible_scene()
    cam = S.small_uniform_camera(96, 80)
    p = gh.RenderParams(spp_side=2, background=BG, method=method, max_points_per_splat=1 << 20,
                        pbvr=dict(radial_correction=1))
    saved = gh.get_tuning()
    try:
        for waves, ppt, block, batch, keep in ((1, 1, 128, 3, True), (2, 4, 256, 0, False)):
            say(f"[owners] n={len(scene)} method={method} waves={waves} ppt={ppt} block={block} batch={batch}")
            gh.set_tuning(particle_waves_per_cu=waves, particles_per_thread=ppt, particle_block=block,
                          batch_ensembles=batch)
            gpu, m = run_pair(scene, None, 0, cam, p, 96, 80, 3, 5, keep=keep)
            assert gpu["stats"].visible_splats > 0
            assert_exact(gpu, m)
    finally:
        gh.set_tuning(**saved)


def test_autotune_keeps_history():
    scene = S.small_uniform(4096, seed=0)
    cam = S.small_uniform_camera(128, 128)
    p = gh.RenderParams(spp_side=2, method=2, max_points_per_splat=16384, pbvr=dict(radial_correction=1))
    saved = gh.get_tuning()
    try:
        r = gh.GpsRenderer(128, 128, 2)
        r.set_gaussians(scene)
        r.set_features(True)
        r.render_ensembles(cam, p, 1, 5)
        before, st_before, sub = r.read_accum(), r.read_stats(), r.submitted()
        mom_before, z_before = r.read_features()
        best, ms = r.autotune(cam, p, trial_ensembles=2)
        say(f"[autotune] {best} {ms:.4f} ms/ens")
        assert np.array_equal(r.read_accum(), before) and r.read_stats() == st_before and r.submitted() == sub
        mom_after, z_after = r.read_features()
        assert np.array_equal(mom_after, mom_before)
        assert np.array_equal(z_after.view(np.uint32), z_before.view(np.uint32))
        assert ms > 0 and best["particles_per_thread"] in (1, 2, 4, 8)
        r.render_ensembles(cam, p, 6, 3)
        after = r.read_accum()
        r.close()
        r2 = gh.GpsRenderer(128, 128, 2)
        r2.set_gaussians(scene)
        r2.render_ensembles(cam, p, 1, 8)
        assert np.array_equal(r2.read_accum(), after)
        r2.close()
    finally:
        gh.set_tuning(**saved)


def test_autotune_sorted_path_searches_digit_widths():
    """On the sorted particle path the autotuner also times the sort's digit widths; the chosen tuning renders
    the same bits as the packed minimum and the history is kept."""
    name, scene = phase_scenes()[3]
    p = gh.RenderParams(spp_side=2, background=BG, method=2, max_points_per_splat=1 << 24, path=2,
                        pbvr=dict(radial_correction=1))
    saved = gh.get_tuning()
    try:
        r = gh.GpsRenderer(W, H, 2)
        r.set_gaussians(to_gpu(scene))
        r.render_ensembles(CAM, p, 3, 1)
        before = r.read_accum()
        best, ms = r.autotune(CAM, p, trial_ensembles=1)
        say(f"[autotune sorted] {best} {ms:.4f} ms/ens")
        assert 4 <= best["sort_digit_bits"] <= 8 and ms > 0
        assert np.array_equal(r.read_accum().view(np.uint32), before.view(np.uint32))
        r.close()
        sorted_vs_atomic(to_gpu(scene), None, 0, CAM, p, W, H, (3, 3))
    finally:
        gh.set_tuning(**saved)


# ------------------------------------------------------------------ batched views in one hipGraph
def test_batch_graph_equals_separate_renders():
    a_scene = S.small_uniform(4096, seed=0)
    b_scene, b_sh = S.clustered(20_000, seed=3, sh_degree=2)
    ra, rb = gh.GpsRenderer(128, 96, 2), gh.GpsRenderer(160, 90, 1)
    ra.set_gaussians(a_scene)
    rb.set_gaussians_sh(b_scene, b_sh, 2)
    cam_a = S.small_uniform_camera(128, 96)
    pa = gh.RenderParams(spp_side=2, method=2, max_points_per_splat=16384, pbvr=dict(radial_correction=1))
    pg = gh.RenderParams(spp_side=2, method=0, max_points_per_splat=4096, path=1)
    pb = gh.RenderParams(spp_side=1, method=1, max_points_per_splat=16384)
    pv = gh.RenderParams(spp_side=2, method=3, max_points_per_splat=16384, pbvr=dict(radial_correction=1))
    po = gh.RenderParams(spp_side=2, method=4, max_points_per_splat=16384)
    # every emitter class of the fused batch: GPS particles (pg), PBVR (pa, pb, pv), screen occupancy (po)
    items = [(ra, cam_a, pa, 11, 5), (rb, S.clustered_camera("mid", 160, 90), pb, 3, 4),
             (ra, cam_a, pg, 1, 3), (rb, S.clustered_camera("far", 160, 90), pb, 7, 70),
             (ra, cam_a, pv, 21, 6), (ra, cam_a, po, 5, 3)]
    expect, expect_st = [], []
    for r, cam, p, seed, count in items:
        r.reset_accum()
        r.reset_stats()
        r.render_ensembles(cam, p, seed, count)
        expect.append(r.read_accum())
        expect_st.append(r.read_stats())
    rb.set_features(True)
    rb.reset_accum()
    rb.render_ensembles(*items[3][1:])
    expect_mom, expect_z = rb.read_features()
    b = gh.Batch()
    saved = gh.get_tuning()
    try:
        # fused (all items in a few launches per round of ensembles) and per item, with and without a graph, with
        # the tuning's round size (one round) and with 16 ensembles per round (5 rounds; items of 3-5 ensembles
        # have none after the first)
        for round_ens, use_graph, per_item in ((0, True, False), (0, False, False), (16, True, False),
                                               (16, False, False), (0, True, True), (0, False, True),
                                               (0, True, False)):
            gh.set_tuning(batch_ensembles=round_ens)
            say(f"[batch] round_ens={round_ens} graph={use_graph} per_item={per_item}")
            b.render(items, use_graph=use_graph, per_item=per_item)
            for i in range(len(items)):
                assert np.array_equal(b.read(i).view(np.uint32), expect[i].view(np.uint32)), i
                assert b.read_stats(i) == expect_st[i], i
            mom, z = rb.read_features()
            assert np.array_equal(mom.view(np.uint32), expect_mom.view(np.uint32))
            assert np.array_equal(z.view(np.uint32), expect_z.view(np.uint32))
            t = b.timing()
            assert t["gpu_ms"] > 0
            if use_graph and per_item:
                assert t["graph_nodes"] > len(items)
            elif use_graph and round_ens == 16:
                # clear, view, then per round: count-scan, (assign, particles) for each of the three emitter
                # classes, resolve
                assert t["graph_nodes"] == 2 + 5 * (2 + 2 * 3), t
    finally:
        gh.set_tuning(**saved)
    # after a batch the renderer carries the last item's history and continues it exactly
    rb.render_ensembles(S.clustered_camera("far", 160, 90), pb, 77, 2)
    rc = gh.GpsRenderer(160, 90, 1)
    rc.set_gaussians_sh(b_scene, b_sh, 2)
    rc.render_ensembles(S.clustered_camera("far", 160, 90), pb, 7, 72)
    assert np.array_equal(rb.read_accum(), rc.read_accum())
    with pytest.raises(gh.GpsError):
        b.render([(ra, cam_a, pa.copy(path=2), 1, 1)])
    b.render([])  # an empty batch is a no-op
    assert b.timing()["graph_nodes"] == 0
    b.close()
    for r in (ra, rb, rc):
        r.close()


def test_fused_batch_many_items_equals_per_item():
    """300 items over six renderers of different sizes and SPP, every method, 0 to 9 ensembles each: the fused
    launches (the particle grid split among up to 300 items on the device) give the bits of each item's own launches,
    with and without a graph, and with 4 ensembles per round."""
    scenes = [S.small_uniform(n, seed=40 + n) for n in (64, 512, 3000)]
    rends = []
    for k, (w, h, spp) in enumerate(((32, 24, 1), (64, 40, 2), (48, 48, 1), (80, 45, 1), (40, 30, 2), (56, 56, 1))):
        r = gh.GpsRenderer(w, h, spp)
        r.set_gaussians(scenes[k % 3])
        rends.append(r)
    rng = np.random.default_rng(5)
    items = []
    for i in range(300):
        r = rends[i % len(rends)]
        w, h = r.size()
        spp = 2 if i % len(rends) in (1, 4) else 1
        m = int(rng.integers(0, 5))
        p = gh.RenderParams(spp_side=spp, method=m, max_points_per_splat=4096, path=1,
                            pbvr=dict(radial_correction=int(rng.integers(0, 2))))
        items.append((r, S.small_uniform_camera(w, h), p, int(rng.integers(1, 1000)), int(rng.integers(0, 10))))
    saved = gh.get_tuning()
    b = gh.Batch()
    try:
        b.render(items, use_graph=False, per_item=True)
        want = [(b.read(i).view(np.uint32).copy(), b.read_stats(i)) for i in range(len(items))]
        for round_ens, use_graph in ((0, True), (4, False), (0, False)):
            gh.set_tuning(batch_ensembles=round_ens)
            b.render(items, use_graph=use_graph)
            for i in range(len(items)):
                assert np.array_equal(b.read(i).view(np.uint32), want[i][0]), (round_ens, use_graph, i)
                assert b.read_stats(i) == want[i][1], (round_ens, use_graph, i)
    finally:
        gh.set_tuning(**saved)
        b.close()
        for r in rends:
            r.close()


C4_SIZES = (20_000, 50_000, 100_000, 250_000, 560_000, 1_000_000, 1_500_000, 2_000_000)
C4_RES = ((640, 360), (1280, 720), (1920, 1080))


def c4_camera(v, w, h, views=16):
    return gh.orbit_camera((0.0, 0.0, 0.0), 6.0, 2.0 * math.pi * v / views, math.radians(15.0), (0.0, 0.0, 1.0),
                           S.FOV_60, w, h)


@pytest.mark.timeout(1200)
def test_config4_batch_graph_equals_separate_renders_and_oracle():
    """Config 4: eight clustered scenes (seeds 10-17; 20k to 2M Gaussians, SH3), each with a 16-view orbit (radius 6,
    elevation 15 degrees) at 640x360 / 1280x720 / 1920x1080 cycling per scene, PBVR Extinction C3+R, 32 ensembles
    per view.  All 128 views in one hipGraph equal the same views rendered one call each (images and counters).
    For every scene one view rendered at 640x360 (2 ensembles) equals the oracle."""
    p = gh.RenderParams(spp_side=1, method=gh.Method.Extinction, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1))
    scenes, rends, items = [], [], []
    for i, n in enumerate(C4_SIZES):
        w, h = C4_RES[i % 3]
        g, sh = S.clustered(n, seed=10 + i, sh_degree=3)
        scenes.append((g, sh))
        r = gh.GpsRenderer(w, h, 1)
        r.set_gaussians_sh(g, sh, 3)
        rends.append(r)
        items += [(r, c4_camera(v, w, h), p, 1 + 1000 * v, 32) for v in range(16)]
    try:
        say(f"[config4] {len(items)} views in one graph")
        b = gh.Batch()
        b.render(items, use_graph=True)
        t = b.timing()
        say(f"[config4] graph: {t['graph_nodes']} nodes, {t['gpu_ms']:.1f} ms")
        for k, (r, cam, q, seed, count) in enumerate(items):
            r.reset_accum()
            r.reset_stats()
            r.render_ensembles(cam, q, seed, count)
            assert np.array_equal(b.read(k).view(np.uint32), r.read_accum().view(np.uint32)), k
            assert b.read_stats(k) == r.read_stats(), k
        b.close()
    finally:
        for r in rends:
            r.close()
    for i, (g, sh) in enumerate(scenes):
        v = (3 * i) % 16
        cam = c4_camera(v, 640, 360)
        say(f"[config4] scene {i} ({C4_SIZES[i]} Gaussians) view {v} at 640x360 against the oracle")
        r = gh.GpsRenderer(640, 360, 1)
        try:
            r.set_gaussians_sh(g, sh, 3)
            r.render_ensembles(cam, p, 1 + 1000 * v, 2)
            acc, st = r.read_accum(), r.read_stats()
        finally:
            r.close()
        m = O.mirror_render(g, sh, 3, ocam(cam), mparams(p), 640, 360, 1 + 1000 * v, 2)
        assert stats_tuple(st) == mirror_stats_tuple(m["stats"]), i
        assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32)), i


# ------------------------------------------------------------------ denoiser features from the resolve
@pytest.mark.parametrize("method,spp,path", [(0, 2, 1), (2, 1, 1), (0, 1, 0)])
def test_features_bit_exact(method, spp, path):
    scene = to_gpu(phase_scenes()[3][1] + phase_scenes()[4][1])
    p = gh.RenderParams(spp_side=spp, method=method, max_points_per_splat=1 << 16, path=path,
                        background=BG, pbvr=dict(radial_correction=1))
    w, h = 40, 32
    r = gh.GpsRenderer(w, h, spp)
    r.set_gaussians(scene)
    r.set_features(True)
    r.render_ensembles(CAM, p, 11, 3)
    r.render_ensembles(CAM, p, 14, 4)
    mom, zmin = r.read_features()
    acc = r.read_accum()
    m_mom = np.zeros((h, w, 4), np.float32)
    m_z = np.full((h, w), np.inf, np.float32)
    m = O.mirror_render(scene, None, 0, ocam(CAM), mparams(p), w, h, 11, 7, features=(m_mom, m_z))
    assert np.array_equal(acc.view(np.uint32), m["accum"].view(np.uint32))
    assert np.array_equal(mom.view(np.uint32), m_mom.view(np.uint32))
    assert np.array_equal(zmin.view(np.uint32), m_z.view(np.uint32))
    r.reset_accum()
    mom, zmin = r.read_features()
    assert not mom.any() and np.isinf(zmin).all()
    r.close()


# ------------------------------------------------------------------ conservative particle cull
@pytest.mark.parametrize("method", [1, 2, 3])
def test_cull_keeps_images_bit_identical(method):
    from gps_hip import denoise as D
    g, sh = S.clustered(60_000, seed=21, sh_degree=2)
    rng = np.random.default_rng(method)
    p = gh.RenderParams(spp_side=1, method=method, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1))
    culled_any = False
    for k in range(4):
        full = D.random_view(rng, 320, 180)
        cam = D.window_camera(full, int(rng.integers(0, 256)), int(rng.integers(0, 116))) if k % 2 else full
        w, h = (64, 64) if k % 2 else (320, 180)
        out = []
        for cull in (False, True):
            r = gh.GpsRenderer(w, h, 1)
            r.set_gaussians_sh(g, sh, 2)
            r.set_cull(cull)
            r.set_debug(gh.DEBUG_KEEP_BUFFERS)
            r.render_ensembles(cam, p, 7 + k, 3)
            out.append((r.read_accum(), r.read_depth_id(), r.read_stats()))
            r.close()
        (a0, d0, s0), (a1, d1, s1) = out
        say(f"[cull m{method} view {k}] visible {s0.visible_splats} -> {s1.visible_splats}, "
            f"candidates {s0.candidates} -> {s1.candidates}")
        assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
        assert np.array_equal(d0, d1)
        assert s1.candidates <= s0.candidates and s1.visible_splats <= s0.visible_splats
        culled_any |= s1.visible_splats < s0.visible_splats
    assert culled_any


# ------------------------------------------------------------------ 3DGS PLY ingest -> renderer
@pytest.mark.parametrize("method,path", [(0, 1), (0, 0), (2, 1), (4, 1)])
def test_ply_fixture_renders_bit_exact(method, path):
    """tests/fixtures/synthetic_sh1.ply through gps_ply_load / gps_ply_to_gpu, framed as gps-verify bench-ply
    frames a file (crates/gps-verify/src/main.rs:107-126: robust sphere, up -Y, yaw 0.6, pitch 0.25)."""
    import os
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "synthetic_sh1.ply")
    g = gh.load_ply(fixture)
    gg = g.to_gpu()
    c, rad = gh.robust_sphere(gg["pos_opacity"][:, :3])
    cam = gh.orbit_camera(tuple(c), gh.frame_sphere_distance(rad), 0.6, 0.25, (0.0, -1.0, 0.0), 0.9, 96, 72)
    p = gh.RenderParams(spp_side=1, method=method, path=path, max_points_per_splat=16384,
                        pbvr=dict(radial_correction=1))
    say(f"[ply] method {method} path {path}")
    gpu, m = run_pair(gg, g.sh_rest, g.sh_degree, cam, p, 96, 72, 1, 3)
    assert_exact(gpu, m)
    assert gpu["stats"].points > 0


# ------------------------------------------------------------------ trajectory export on one batch graph
def test_trajectory_export_matches_single_renders():
    """Each exported frame (PNG and CSV counters) equals a separate render of its camera."""
    import io
    import os
    import zipfile
    from gps_hip import trajectory as T
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "synthetic_sh1.ply")
    g = gh.load_ply(fixture)
    gg = g.to_gpu()
    c, rad = gh.robust_sphere(g.position)
    key = {"target": [float(x) for x in c], "distance": gh.frame_sphere_distance(rad), "yaw": 0.6, "pitch": 0.25,
           "fov_y": 0.9, "up_y": -1}
    p = gh.RenderParams(spp_side=1, method=2, max_points_per_splat=16384, pbvr=dict(radial_correction=1))
    r = gh.GpsRenderer(80, 60, 1)
    r.set_gaussians_sh(gg, g.sh_rest, g.sh_degree)
    keys = T.orbit_keys(key)
    data, rows, t = T.render_trajectory(r, p, keys, 5, 4, "pbvr-ext-c3r", first_seed=3)
    z = zipfile.ZipFile(io.BytesIO(data))
    names = sorted(z.namelist())
    assert names == [f"frame_{i:04d}.png" for i in range(5)] + ["session.json", "stats.csv"]
    csv = z.read("stats.csv").decode().strip().split("\n")
    assert csv[0].split(",") == ["frame"] + T.CSV_COLS and len(csv) == 6
    for i in range(5):
        k = T.cam_at(keys, i / 4)
        r.reset_accum()
        r.reset_stats()
        cam = T.key_camera(k, 80, 60)
        r.render_ensembles(cam, p, 3, 4)
        want = T.png_rgba8(r.composite(p))
        assert z.read(f"frame_{i:04d}.png") == want, i
        st = r.read_stats()
        cols = dict(zip(["frame"] + T.CSV_COLS, csv[1 + i].split(",")))
        assert int(cols["points"]) == st.points and int(cols["candidates"]) == st.candidates
        assert int(cols["accumulated"]) == 4 and int(cols["visible"]) == st.visible_splats
        # the frame against the oracle: its accumulation, counters and composited PNG
        m = O.mirror_render(gg, g.sh_rest, g.sh_degree, ocam(cam), mparams(p), 80, 60, 3, 4)
        assert np.array_equal(r.read_accum().view(np.uint32), m["accum"].view(np.uint32)), i
        assert stats_tuple(st) == mirror_stats_tuple(m["stats"]), i
        png = T.png_rgba8(O.mirror_composite(m["accum"], p.exposure, p.srgb_output, 80, 60))
        assert z.read(f"frame_{i:04d}.png") == png, i
    assert t["graph_nodes"] > 5
    r.close()


# ------------------------------------------------------------------ count-only path for off-screen splats
def _render_with(count_only, scene, sh, deg, cam, p, w, h, seed, count, peaks=False):
    old = gh.get_tuning()
    gh.set_tuning(count_only=count_only)
    try:
        r = gh.GpsRenderer(w, h, p.spp_side)
        r.set_gaussians_sh(scene, sh, deg) if sh is not None else r.set_gaussians(scene)
        r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.render_ensembles(cam, p, seed, count)
        out = (r.read_accum(), r.read_depth_id(), r.read_stats())
        if peaks:
            out = out + (r.read_queue_peaks(),)
        r.close()
    finally:
        gh.set_tuning(**old)
    return out


@pytest.mark.parametrize("case", [(2, 3, 1, 0, 1), (2, 3, 1, 0, 0), (2, 3, 0, 0, 1), (1, 0, 1, 0, 0), (3, 0, 1, 1, 1),
                                  (2, 1, 1, 0, 1)])
def test_count_only_matches_mirror_in_windows(case):
    """Windows of random views see a few percent of the splats: most particles go through the count-only path.
    Every counter, the depth|id buffer and the image equal the oracle's."""
    from gps_hip import denoise as D
    method, cal, radial, centre, jitter = case
    g, sh = S.clustered(20_001, seed=31, sh_degree=1)   # odd: the splat order's counter word sits after a pad
    rng = np.random.default_rng(17 + method)
    p = gh.RenderParams(spp_side=1, method=method, max_points_per_splat=4096,
                        pbvr=dict(calibration=cal, radial_correction=radial, centre_depth=centre,
                                  low_pass_jitter=jitter))
    for k in range(3):
        cam = D.window_camera(D.random_view(rng, 320, 180), int(rng.integers(0, 256)), int(rng.integers(0, 116)))
        say(f"[count-only vs mirror] case={case} view {k}")
        r = gh.GpsRenderer(64, 64, 1)
        r.set_gaussians_sh(g, sh, 1)
        r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.render_ensembles(cam, p, 5 + k, 2)
        gpu = dict(accum=r.read_accum(), stats=r.read_stats(), bufs=None)
        r.close()
        m = O.mirror_render(g, sh, 1, ocam(cam), mparams(p), 64, 64, 5 + k, 2)
        assert_exact(gpu, m, check_buffers=False)


@pytest.mark.parametrize("view", ["near", "mid", "far"])
@pytest.mark.parametrize("method,radial,jitter", [(2, 1, 1), (2, 1, 0), (2, 0, 1), (1, 1, 1), (3, 1, 1)])
def test_count_only_equals_full_path_560k(view, method, radial, jitter):
    """The 560k-splat clustered scene at 1280x720: with and without the count-only path, the image, the
    depth|id buffers and every counter are identical (near: about 95% of the particles are off-screen)."""
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    cam = S.clustered_camera(view)
    p = gh.RenderParams(spp_side=1, method=method, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=radial, low_pass_jitter=jitter))
    a0, d0, s0 = _render_with(0, g, sh, 3, cam, p, 1280, 720, 11, 3)
    a1, d1, s1 = _render_with(1, g, sh, 3, cam, p, 1280, 720, 11, 3)
    say(f"[count-only 560k {view} m{method} r{radial} j{jitter}] points {s0.points} / {s1.points}")
    assert stats_tuple(s0) == stats_tuple(s1)
    assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
    assert np.array_equal(d0, d1)


@pytest.mark.parametrize("cap_log2", [26, 24])
def test_count_only_above_24_bit_cap(cap_log2):
    """A splat beside the camera (every particle off-screen, so on the count-only path) with about 20M particles:
    particle indices beyond 2^24 within one splat (its entries carry the particle's seed, not its index), next to an
    on-screen splat.  At cap 2^26 nothing is truncated; at cap 2^24 the big splat is.  Image, counters, per-splat
    counts and offsets equal the oracle's, and the count-only buffers were used."""
    scene = to_gpu([dict(pos=(10.0, 0.0, 3.0), log_scale=(math.log(0.5),) * 3, opacity=0.6, color=(0.2, 0.5, 0.9),
                         rot_xyzw=(0.0, 0.0, 0.0, 1.0)), phase_scenes()[1][1][0]])
    p = gh.RenderParams(spp_side=1, method=2, max_points_per_splat=1 << cap_log2, density_scale=900.0,
                        background=BG, pbvr=dict(calibration=3, radial_correction=1, low_pass_jitter=1))
    old = gh.get_tuning()
    gh.set_tuning(count_only=1)
    try:
        r = gh.GpsRenderer(W, H, 1)
        r.set_gaussians(scene)
        r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.render_ensembles(CAM, p, 3, 1)
        gpu = dict(accum=r.read_accum(), stats=r.read_stats(), bufs=r.read_counts())
        gpu["bufs"]["depth_id"] = r.read_depth_id()
        pk = r.read_queue_peaks()
        r.close()
    finally:
        gh.set_tuning(**old)
    st = gpu["stats"]
    say(f"[count-only cap 2^{cap_log2}] counts {gpu['bufs']['counts']}, points {st.points}, truncated "
        f"{st.truncated_splats}, work peak {pk[2]} / {pk[3]}")
    m = O.mirror_render(scene, None, 0, ocam(CAM), mparams(p), W, H, 3, 1, want_buffers=True)
    assert_exact(gpu, m)
    assert pk[2] > 0 and pk[2] < pk[3] and pk[0] <= pk[1]
    if cap_log2 == 26:
        assert gpu["bufs"]["counts"][0] > (1 << 24) and st.truncated_splats == 0
    else:
        assert gpu["bufs"]["counts"][0] == 1 << 24 and st.truncated_splats == 1


def test_count_only_last_chunk_flush_keeps_remainders():
    """Regression: after a block's last chunk flushes the emit_pbvr queue, the per-wave work buffers' remainders
    must be appended after the queue counter's reset.  Few resident blocks (4 waves per CU) with one particle
    per thread give every block many chunks and many queue flushes; over several seeds the mid view's counters
    and images must equal those of the path turned off."""
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    cam = S.clustered_camera("mid", 640, 360)
    p = gh.RenderParams(spp_side=1, method=2, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1, low_pass_jitter=1))
    saved = gh.get_tuning()
    try:
        for seed in range(101, 109):
            gh.set_tuning(particle_waves_per_cu=4, particles_per_thread=1, particle_block=256)
            a0, _, s0 = _render_with(0, g, sh, 3, cam, p, 640, 360, seed, 2)
            a1, _, s1, pk = _render_with(1, g, sh, 3, cam, p, 640, 360, seed, 2, peaks=True)
            say(f"[count-only last-chunk flush] seed {seed}: points {s0.points} / {s1.points}, queue peak "
                f"{pk[0]} / {pk[1]}, work peak {pk[2]} / {pk[3]}")
            assert pk[0] <= pk[1] and pk[2] < pk[3] and pk[0] > 0
            assert stats_tuple(s0) == stats_tuple(s1)
            assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
    finally:
        gh.set_tuning(**saved)


@pytest.mark.parametrize("view", ["near", "mid"])
def test_count_only_under_tunings(view):
    """The keep-draw shortcut and the per-wave count_only buffer (64 x (PPT + 1) entries) under every particles-
    per-thread value and workgroup size, with and without the splat order: identical to the path turned off."""
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    cam = S.clustered_camera(view)
    p = gh.RenderParams(spp_side=1, method=2, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1, low_pass_jitter=1))
    a0, d0, s0 = _render_with(0, g, sh, 3, cam, p, 1280, 720, 21, 2)
    saved = gh.get_tuning()
    try:
        for waves, ppt, block, keep in ((8, 1, 128, False), (16, 2, 512, True), (64, 8, 256, False),
                                        (256, 4, 256, True), (32, 4, 128, False)):
            say(f"[count-only tunings {view}] waves={waves} ppt={ppt} block={block} keep={keep}")
            gh.set_tuning(particle_waves_per_cu=waves, particles_per_thread=ppt, particle_block=block, count_only=1)
            r = gh.GpsRenderer(1280, 720, 1)
            r.set_gaussians_sh(g, sh, 3)
            if keep:
                r.set_debug(gh.DEBUG_KEEP_BUFFERS)
            r.render_ensembles(cam, p, 21, 2)
            a1, s1 = r.read_accum(), r.read_stats()
            d1 = r.read_depth_id() if keep else None
            pk = r.read_queue_peaks()
            r.close()
            say(f"    queue peak {pk[0]} / {pk[1]}, work peak {pk[2]} / {pk[3]}")
            assert pk[0] <= pk[1] and pk[2] < pk[3] and pk[2] > 0
            assert stats_tuple(s1) == stats_tuple(s0)
            assert np.array_equal(a1.view(np.uint32), a0.view(np.uint32))
            if keep:
                assert np.array_equal(d1, d0)
    finally:
        gh.set_tuning(**saved)


def test_async_stats_readback():
    """gps.rs:1005-1038: request_stats starts a non-blocking readback (one in flight), latest_stats delivers the
    last finished one with its issue number, stats_issued counts the readbacks; a snapshot whose issue is above the
    count read right after some work was enqueued shows that work."""
    p = gh.RenderParams(spp_side=2, background=BG, method=2, pbvr=dict(radial_correction=1))
    r = gh.GpsRenderer(W, H, 2)
    try:
        r.set_gaussians(to_gpu(phase_scenes()[3][1]))
        assert r.stats_issued() == 0 and r.latest_stats().issue == 0 and r.latest_stats().points == 0
        for k in range(3):
            r.render_ensembles(CAM, p, 1 + 4 * k, 4)
            before = r.stats_issued()
            r.request_stats()
            r.request_stats()             # no-op while the first is in flight, or a second readback
            assert before + 1 <= r.stats_issued() <= before + 2
            r.wait_idle()
            s = r.latest_stats()
            ref = r.read_stats()
            say(f"[async stats] round {k}: issue {s.issue} of {r.stats_issued()}, points {s.points}")
            assert s.issue > before and s.issue <= r.stats_issued()
            assert stats_tuple(s) == stats_tuple(ref) and s.ensembles == 4 * (k + 1)
    finally:
        r.close()
