"""The f32 mirror (oracle/gps_mirror.c) against the f64 oracle and libm.

The mirror restates shaders/gps_splat.wgsl in f32; the reference gates its GPU
(WGSL) path against the f64 CPU oracle with fixed thresholds
(crates/gps-render/src/verify.rs:120-123, 384-390, 720-722).  The same gates,
run here on the CPU, pin the mirror, which is the bit-exact oracle of the HIP
kernels.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from scenes import BG, crowd, iso, many_splats, phase_scenes, to_gpu, to_image

W = H = 96


def params(**kw):
    p = O.MParams()
    p.spp_side = 2
    p.density_scale = 1.0
    p.max_points_per_splat = 1 << 16
    p.near = 0.05
    p.background[:] = BG
    p.exposure = 1.0
    p.srgb_output = 1
    p.exact_radius = 0
    p.sh_degree = 3
    p.path = 1
    p.method = 0
    p.pbvr.calibration = 3
    p.pbvr.radial_correction = 0
    p.pbvr.centre_depth = 0
    p.pbvr.low_pass_jitter = 1
    p.pbvr.base_k = 512.0
    p.pbvr.reference_pixel_length = 0.01
    for k, v in kw.items():
        if k.startswith("pbvr_"):
            setattr(p.pbvr, k[5:], v)
        else:
            setattr(p, k, v)
    return p


CAM = O.default_camera(200.0)


def ulp_err(y, ref):
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(y.astype(np.float64) - ref) / sp


@pytest.mark.parametrize("name,fn,lo,hi", [("expf", np.exp, -103.0, 88.7), ("logf", np.log, 1e-30, 1e30),
                                           ("sinf", np.sin, 0.0, 6.2832), ("cosf", np.cos, 0.0, 6.2832)])
def test_f32_elementary_functions_within_2ulp(name, fn, lo, hi):
    rng = np.random.default_rng(7)
    if name == "logf":
        x = np.exp(rng.uniform(math.log(lo), math.log(hi), 400_000)).astype(np.float32)
        x = np.concatenate([x, rng.uniform(0.5, 2.0, 200_000).astype(np.float32)])
    else:
        x = rng.uniform(lo, hi, 400_000).astype(np.float32)
    y = O.mirror_fn("mir_" + name + "_n", x)
    ref = fn(x.astype(np.float64))
    m = np.abs(ref) > 1e-3
    assert ulp_err(y[m], ref[m]).max() <= 2.0
    if name in ("sinf", "cosf"):
        assert np.abs(y.astype(np.float64) - ref).max() <= 2.0 * 2.0 ** -24


def test_f32_elementary_special_values():
    inf = np.float32(np.inf)
    e = O.mirror_fn("mir_expf_n", np.array([0.0, 89.0, -104.0, -np.inf], np.float32))
    assert e[0] == 1.0 and e[1] == inf and e[2] == 0.0 and e[3] == 0.0
    l = O.mirror_fn("mir_logf_n", np.array([1.0, 0.0, np.inf, 1e-45, 2.0], np.float32))
    assert l[0] == 0.0 and l[1] == -inf and l[2] == inf and abs(l[3] - math.log(float(np.float32(1e-45)))) < 1e-4
    assert abs(l[4] - math.log(2.0)) < 1e-7
    assert np.isnan(O.mirror_fn("mir_logf_n", np.array([-1.0], np.float32))[0])


def test_mirror_dilog_and_radius_track_f64():
    for x in np.linspace(0.001, 0.999, 101):
        assert abs(O.lib.mir_dilog(float(x)) - O.lib.orc_dilog(float(x))) < 2e-6
    lut = O.cached_lut()
    lp = lut.ctypes.data_as(O.P(O.f32))
    for o in (0.05, 0.3, 0.7, 0.95):
        for u in (0.01, 0.3, 0.6, 0.9, 0.999):
            ex = O.lib.orc_sample_corrected_radius(o, u)
            assert abs(O.lib.mir_corrected_radius(o, u) - ex) < 2e-3 * max(1.0, ex)
            assert abs(O.lib.mir_corrected_radius_lut(lp, o, u) - ex) < 0.011
            assert abs(O.lib.mir_radial_keep(o, u) - O.lib.orc_radial_keep_probability(o, u)) < 1e-5


def test_mirror_phase0_two_pass():
    """lib.rs:99-192 CPU reference restated in numpy (ties forced by the 16-bit key mask)."""
    PIX, SAMPLES = 4096, 200_000
    for e in range(3):
        seed = 1000 + e * 7
        d = np.zeros(PIX, np.uint32)
        c = np.zeros(PIX, np.uint32)
        O.lib.mir_phase0(seed, 0, PIX, SAMPLES, O.ptr(d, O.u32), O.ptr(c, O.u32))
        i = np.arange(SAMPLES, dtype=np.uint64)
        m = np.uint64(0xFFFFFFFF)

        def pcg(v):
            s = (v * np.uint64(747796405) + np.uint64(2891336453)) & m
            w = (((s >> ((s >> np.uint64(28)) + np.uint64(4))) ^ s) * np.uint64(277803737)) & m
            return (w >> np.uint64(22)) ^ w

        h0 = pcg(i ^ np.uint64(O.lib.mir_pcg(seed)))
        h1 = pcg(h0)
        h2 = pcg(h1)
        pix = (h0 % np.uint64(PIX)).astype(np.int64)
        key = (((np.uint64(0x3F800000) | (h1 & np.uint64(0x7FFFFF))) ^ np.uint64(0x80000000)) & np.uint64(0xFFFF0000))
        rd = np.full(PIX, 0xFFFFFFFF, np.uint64)
        np.minimum.at(rd, pix, key)
        win = rd[pix] == key
        rc = np.full(PIX, 0xFFFFFFFF, np.uint64)
        np.minimum.at(rc, pix[win], (h2 & np.uint64(0xFFFFFF))[win])
        assert (d == rd).all() and (c == rc).all()
        assert win.sum() > PIX  # ties actually occur


def mirror_mean(scene, p, sets, seed0=1, cam=CAM, w=W, h=H):
    sp = to_gpu(scene)
    r = O.mirror_render(sp, None, 0, cam, p, w, h, seed0, sets)
    return to_image(r["accum"]), r["stats"]


@pytest.mark.parametrize("idx", range(5))
def test_phase2_gps_mirror_vs_oracle(idx):
    """verify.rs:92-144: both paths vs the CPU Monte Carlo oracle and the analytic image."""
    name, scene = phase_scenes()[idx]
    ana = O.render_analytic(scene, CAM, W, H, BG)
    seen = []
    for path in (0, 1):
        p = params(path=path)
        img, st = mirror_mean(scene, p, 300)
        opt = O.default_mc_options(per_particle_seed=path, max_points_per_splat=p.max_points_per_splat)
        cpu, cst = O.render_monte_carlo(scene, CAM, W, H, 2, 300, BG, 1, opt)
        assert O.psnr(img, cpu) >= 60.0, name
        assert abs(float(st[0]) - cst.points) <= 4.0 * max(math.sqrt(2.0 * cst.points), 1.0)
        assert st[4] == 0
        assert O.psnr(img, ana) >= 35.0 and O.ssim(img, ana) >= 0.98 and st[1] == 0
        seen.append(st.copy())
    assert (seen[0] == seen[1]).all()


def test_phase2_cap_truncation_counters_agree():
    """verify.rs:187-191"""
    scene = phase_scenes()[2][1]
    seen = []
    for path in (0, 1):
        _, st = mirror_mean(scene, params(path=path, max_points_per_splat=50), 8)
        seen.append(st.copy())
    assert (seen[0] == seen[1]).all() and seen[0][1] > 0 and seen[0][2] > 0


PBVR_CASES = [
    ("Proportional C0", 1, 0, 0, 0, 1), ("Proportional C3", 1, 3, 0, 0, 1), ("Extinction C0", 2, 0, 0, 0, 1),
    ("Extinction C1", 2, 1, 0, 0, 1), ("Extinction C2", 2, 2, 0, 0, 1), ("Extinction C3", 2, 3, 0, 0, 1),
    ("Extinction C3 no-jitter", 2, 3, 0, 0, 0), ("Extinction C3 centre-depth", 2, 3, 0, 1, 1),
    ("Extinction C3+R", 2, 3, 1, 0, 1), ("Extinction C3+R centre-depth", 2, 3, 1, 1, 1),
]


def pbvr_params(method, cal, radial, centre, jitter, **kw):
    return params(method=method, pbvr_calibration=cal, pbvr_radial_correction=radial, pbvr_centre_depth=centre,
                  pbvr_low_pass_jitter=jitter, max_points_per_splat=1 << 24, **kw)


@pytest.mark.parametrize("si", [1, 3, 4])
@pytest.mark.parametrize("case", PBVR_CASES, ids=[c[0] for c in PBVR_CASES])
def test_phase3_pbvr_mirror_vs_oracle(si, case):
    """verify.rs:368-403 + 429-464: PSNR >= 50 dB and counts within 4 sigma; C3+R >= 30 dB vs analytic."""
    _, method, cal, radial, centre, jitter = case
    scene = phase_scenes()[si][1]
    img, st = mirror_mean(scene, pbvr_params(method, cal, radial, centre, jitter), 120)
    opt = O.default_particle_options(level=cal, rule=0 if method == 1 else 1, radial_correction=radial,
                                     centre_depth=centre, low_pass_jitter=jitter, per_particle_seed=1)
    cpu, cst = O.render_particles_3d(scene, CAM, W, H, 2, 120, BG, 1, opt)
    assert O.psnr(img, cpu) >= 50.0
    assert abs(float(st[0]) - cst.points) <= 4.0 * max(math.sqrt(2.0 * cst.points), 1.0)
    assert abs(float(st[5]) - cst.candidates) <= 4.0 * max(math.sqrt(2.0 * cst.candidates), 1.0)
    assert st[4] == 0 and st[1] == 0
    if radial:
        assert O.psnr(img, O.render_analytic(scene, CAM, W, H, BG)) >= 30.0


@pytest.mark.parametrize("si", [1, 3, 4])
def test_phase3_occupancy_and_view_conditioned(si):
    """verify.rs:467-516"""
    scene = phase_scenes()[si][1]
    ana = O.render_analytic(scene, CAM, W, H, BG)
    img, st = mirror_mean(scene, params(method=4, max_points_per_splat=1), 120)
    assert O.psnr(img, ana) >= 30.0 and st[4] == 0 and st[1] == 0
    assert st[5] <= len(scene) * W * H * 4 * 120
    vc = pbvr_params(3, 0, 0, 0, 1, pbvr_base_k=4096.0)
    img, st_vc = mirror_mean(scene, vc, 120)
    _, st_gps = mirror_mean(scene, params(max_points_per_splat=1 << 24), 120)
    assert O.psnr(img, ana) >= 30.0
    assert abs(float(st_vc[0]) - float(st_gps[0])) <= 4.0 * math.sqrt(2.0 * float(st_gps[0]))
    assert st_vc[6] == 0


def test_phase3_view_conditioned_undercovered_reported():
    """verify.rs:532-539"""
    _, st = mirror_mean(phase_scenes()[1][1], pbvr_params(3, 0, 0, 0, 1), 8)
    assert st[6] > 0


def test_phase4_sh_degrees_vs_eval_sh():
    """verify.rs:656-732: each SH degree vs the f64 eval_sh colours through the Monte Carlo oracle (>= 60 dB),
    and each degree must change the image (< 45 dB between consecutive degrees)."""
    stride = 15
    seed = 99

    def rnd():
        nonlocal seed
        from scenes import pcg_py
        seed = pcg_py(seed)
        return seed / 4294967296.0 - 0.5

    dcs, rests, base = [], [], []
    for i in range(4):
        dc = np.array([rnd(), rnd(), rnd()]) * 1.2
        dcs.append(dc)
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
        r = O.mirror_render(sp, rest32, 3, CAM, params(sh_degree=deg), W, H, 1, 300)
        gpu = to_image(r["accum"])
        cpu, _ = O.render_monte_carlo(expected, CAM, W, H, 2, 300, BG, 1, O.default_mc_options(per_particle_seed=1))
        assert O.psnr(gpu, cpu) >= 60.0, deg
        images.append(gpu)
    for d in range(1, 4):
        assert O.psnr(images[d], images[d - 1]) < 45.0


def test_overflow_skips_the_ensemble():
    """gps_splat.wgsl:636-639, 650-654: a particle total above 2^32-1 skips the ensemble (counted, accum untouched)."""
    scene = phase_scenes()[3][1]
    p = params(density_scale=1.0e7, max_points_per_splat=0xFFFFFFFF, path=1)
    r = O.mirror_render(to_gpu(scene), None, 0, CAM, p, 16, 16, 1, 1, want_buffers=True)
    assert r["skipped"] == 1 and r["total"] == 0 and r["stats"][7] == 1
    assert (r["accum"] == 0).all()
    assert r["stats"][0] > 0xFFFFFFFF  # the Poisson counts were still counted



@pytest.mark.parametrize("method,spp", [(0, 2), (2, 1)])
def test_mirror_features_match_per_ensemble_restatement(method, spp):
    """Denoiser features of a 5-ensemble call equal those rebuilt from five 1-ensemble calls: each one's image
    (a fresh accumulation holds exactly one ensemble's values) squared and summed in ensemble order in f32,
    covered subpixels counted from its depth|id buffer, nearest depth from the keys' low 31 bits."""
    scene = to_gpu(phase_scenes()[3][1] + phase_scenes()[4][1])
    p = params(method=method, spp_side=spp, pbvr_radial_correction=1)
    w, h = 40, 32
    mom = np.zeros((h, w, 4), np.float32)
    zmin = np.full((h, w), np.inf, np.float32)
    r = O.mirror_render(scene, None, 0, CAM, p, w, h, 11, 5, features=(mom, zmin))
    ex_m = np.zeros((h, w, 4), np.float32)
    ex_z = np.full((h, w), 0x7F800000, np.uint32)
    acc = np.zeros((h, w, 4), np.float32)
    for e in range(5):
        one = O.mirror_render(scene, None, 0, CAM, p, w, h, 11 + e, 1, want_buffers=True)
        v = one["accum"][..., :3]
        acc[..., :3] = acc[..., :3] + v
        ex_m[..., :3] = ex_m[..., :3] + v * v
        d = one["depth_id"].reshape(h, w, spp * spp)
        hit = d != np.uint64(~np.uint64(0))
        ex_m[..., 3] = ex_m[..., 3] + hit.sum(-1).astype(np.float32)
        zb = np.where(hit, (d >> np.uint64(32)).astype(np.uint32) & np.uint32(0x7FFFFFFF), np.uint32(0x7F800000))
        ex_z = np.minimum(ex_z, zb.min(-1))
    assert np.array_equal(r["accum"][..., :3], acc[..., :3])
    assert np.array_equal(mom.view(np.uint32), ex_m.view(np.uint32))
    assert np.array_equal(zmin.view(np.uint32), ex_z)
    assert ex_m[..., 3].max() > 0 and np.isfinite(zmin).any() and np.isinf(zmin).any()
