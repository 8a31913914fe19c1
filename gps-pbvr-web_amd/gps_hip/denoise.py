"""Training and evaluation of the learned denoiser on pairs the renderer generates.

A pair is the feature image of an `e`-ensemble PBVR Extinction C3+R render (input) and the mean of an
independent many-ensemble plain render of the same view (target).  Training views are 128 x 128 windows of
1280 x 720 orbit views of clustered scenes: a window is rendered by its own camera with the principal point
shifted, which renders the particles of that window of the full image at a fraction of the cost.  Inputs
are rendered at several ensemble counts per target, so one network serves every count (the standard error
feature tells it the noise level).

    python -m gps_hip.denoise --out weights.f32      # renders pairs, trains, saves the parameter vector

PSNR follows crates/gps-core/src/oracle.rs:503-514 (linear RGB, MSE over all channels, peak 1, capped at
120 dB).
"""
import argparse
import ctypes
import math
import os
import sys
import time

import numpy as np

from . import (Denoiser, GpsCamera, GpsRenderer, Method, RenderParams, orbit_camera)
from . import scenes as S

FULL_W, FULL_H = 1280, 720
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "denoiser_ext_c3r.f32")


def params_ext_c3r():
    return RenderParams(spp_side=1, method=Method.Extinction, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1))


def psnr(a, b):
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2
    mse = float(d.mean())
    return 120.0 if mse <= 1e-20 else 10.0 * math.log10(1.0 / mse)


def mean_image(acc):
    return acc[..., :3] / np.maximum(acc[..., 3:4], 1.0)


def random_view(rng, w=FULL_W, h=FULL_H, dist_range=(2.0, 15.0)):
    """Orbit camera around the origin: distance log-uniform in dist_range, yaw uniform, pitch in [-0.35, 0.7]."""
    dist = math.exp(rng.uniform(math.log(dist_range[0]), math.log(dist_range[1])))
    return orbit_camera((0.0, 0.0, 0.0), dist, rng.uniform(0.0, 2.0 * math.pi), rng.uniform(-0.35, 0.7),
                        (0.0, 0.0, 1.0), S.FOV_60, w, h)


def window_camera(cam, x0, y0):
    """Camera that renders the window starting at pixel (x0, y0) of `cam`'s image."""
    c = GpsCamera()
    ctypes.memmove(ctypes.byref(c), ctypes.byref(cam), ctypes.sizeof(GpsCamera))
    c.cx = cam.cx - x0
    c.cy = cam.cy - y0
    if c.cx == 0.0 and c.cy == 0.0:  # (0, 0) means "image centre" to the renderer
        c.cx = 1e-9
    return c


def build_pairs(d, seeds, targets, crop=128, in_ens=(4, 16, 64, 256, 768), tgt_ens=2048, n_splats=560_000, rng_seed=0,
                log=print, dist_range=(2.0, 15.0), time_budget=None):
    """Renders `targets` target windows spread over the scenes `seeds` and, per target, one input per count in
    `in_ens`; adds every (input, target) pair to the denoiser.  crop=0: full 1280 x 720 frames instead of windows.
    With a time budget (seconds) it stops adding targets once the budget is spent.  Returns the targets rendered."""
    rng = np.random.default_rng(rng_seed)
    p = params_ext_c3r()
    w, h = (crop, crop) if crop else (FULL_W, FULL_H)
    inp = GpsRenderer(w, h, 1)
    tgt = GpsRenderer(w, h, 1)
    inp.set_features(True)
    inp.set_cull(True)  # a window sees a few percent of the splats; the cull keeps the image bit for bit
    tgt.set_cull(True)
    per_scene = [targets // len(seeds) + (1 if i < targets % len(seeds) else 0) for i in range(len(seeds))]
    k = 0
    t0 = time.perf_counter()
    for si, seed in enumerate(seeds):
        if per_scene[si] == 0:
            continue
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            log(f"[pairs] time budget of {time_budget:.0f} s spent after {k} targets")
            break
        g, sh = S.clustered(n_splats, seed=seed, sh_degree=3)
        inp.set_gaussians_sh(g, sh, 3)
        tgt.set_gaussians_sh(g, sh, 3)
        for _ in range(per_scene[si]):
            full = random_view(rng, dist_range=dist_range)
            cam = window_camera(full, int(rng.integers(0, FULL_W - crop + 1)),
                                int(rng.integers(0, FULL_H - crop + 1))) if crop else full
            tgt.reset_accum()
            tgt.render_ensembles(cam, p, 1_000_000 + 4096 * k, tgt_ens)
            for j, e in enumerate(in_ens):
                inp.reset_accum()
                inp.render_ensembles(cam, p, 1 + 4096 * k + 512 * j, e)
                d.add_pair(inp, tgt)
            k += 1
        log(f"[pairs] scene {seed}: {k}/{targets} targets, {d.pairs()} pairs, {time.perf_counter() - t0:.1f} s")
    inp.close()
    tgt.close()
    return k


def train(d, steps=16000, batch=16, lr=1e-3, seed=1, chunks=8, log=print):
    """Adam with the learning rate halved every steps/chunks steps after the first half."""
    losses = []
    per = max(1, steps // chunks)
    for c in range(chunks):
        rate = lr if c < chunks // 2 else lr * 0.5 ** (c - chunks // 2 + 1)
        l = d.train(per, batch, lr=rate, seed=seed + c)
        losses.append(l)
        log(f"[train] chunk {c + 1}/{chunks} lr {rate:.2e} loss {l[:8].mean():.5g} -> {l[-8:].mean():.5g}")
    return np.concatenate(losses)


def ladder(lo=1, hi=4096, steps_per_octave=8):
    """Ensemble counts lo..hi in geometric steps of 2^(1/steps_per_octave) (about 9%), rounded, distinct."""
    out, k = [], 0
    while True:
        e = int(round(2.0 ** (k / steps_per_octave)))
        k += 1
        if e < lo or (out and e == out[-1]):
            continue
        if e > hi:
            break
        out.append(e)
    return tuple(out)


def denoised_time_to_psnr(r, d, cam, p, ms_per_ens, ref, target_db=40.0, ladder_=None, fallback_ratio=4.0,
                          first_seed=1):
    """Smallest ensemble count on the ladder whose denoised mean reaches target_db against `ref`, and its GPU
    time: the ensembles' render time plus the denoiser's measured time.  `r` must have its features enabled."""
    ladder_ = ladder_ or ladder(4)
    r.reset_accum()
    done, curve = 0, []
    for e in ladder_:
        r.render_ensembles(cam, p, first_seed + done, e - done)
        done = e
        info = d.apply(r, fallback_ratio=fallback_ratio)
        q = psnr(d.read()[..., :3], ref)
        curve.append((e, round(q, 2), round(info["fallback_rate"], 3)))
        if q >= target_db:
            return {"ensembles": e, "gpu_ms": round(e * ms_per_ens + info["gpu_ms"], 3), "psnr_db": round(q, 2),
                    "denoise_ms": round(info["gpu_ms"], 3), "fallback_rate": round(info["fallback_rate"], 4),
                    "curve": curve}
    return {"ensembles": None, "gpu_ms": None, "psnr_db": curve[-1][1], "curve": curve}


def evaluate_heldout(d, seeds=range(200, 210), in_ens=4, ref_ens=1024, n_splats=560_000, fallback_ratio=4.0,
                     rng_seed=12345, log=print):
    """Per held-out scene, one random 1280 x 720 view: PSNR and max abs error of the plain and the denoised
    in_ens-ensemble image against the plain ref_ens-ensemble render, and the fallback rate."""
    rng = np.random.default_rng(rng_seed)
    p = params_ext_c3r()
    r = GpsRenderer(FULL_W, FULL_H, 1)
    r.set_features(True)
    r.set_cull(True)  # the same images, fewer wasted particles
    rows = []
    for seed in seeds:
        g, sh = S.clustered(n_splats, seed=seed, sh_degree=3)
        r.set_gaussians_sh(g, sh, 3)
        cam = random_view(rng)
        r.reset_accum()
        r.render_ensembles(cam, p, 1_000_000, ref_ens)
        ref = mean_image(r.read_accum())
        r.reset_accum()
        r.render_ensembles(cam, p, 1, in_ens)
        plain = mean_image(r.read_accum())
        info = d.apply(r, fallback_ratio=fallback_ratio)
        den = d.read()[..., :3]
        row = {"seed": seed, "psnr_plain": round(psnr(plain, ref), 3), "psnr_denoised": round(psnr(den, ref), 3),
               "max_abs_plain": round(float(np.abs(plain - ref).max()), 4),
               "max_abs_denoised": round(float(np.abs(den - ref).max()), 4),
               "fallback_rate": round(info["fallback_rate"], 4), "denoise_ms": round(info["gpu_ms"], 3)}
        rows.append(row)
        log(f"[heldout] {row}")
    r.close()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--seeds", default="100-139", help="training scenes, a-b inclusive")
    ap.add_argument("--targets", type=int, default=1200)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--in-ens", default="4,16,64,256,768")
    ap.add_argument("--tgt-ens", type=int, default=2048)
    ap.add_argument("--splats", type=int, default=560_000)
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default=None, help="start from this parameter file")
    ap.add_argument("--dist", default="2,15", help="orbit distance range of the training views, lo,hi")
    ap.add_argument("--rng-seed", type=int, default=0, help="seed of the view and window draws")
    ap.add_argument("--time-budget", type=float, default=None, help="seconds of pair rendering at most")
    ap.add_argument("--config5", action="store_true",
                    help="the denoiser workload as specified: 2000 full-frame 1280x720 pairs of 4-ensemble inputs "
                         "and 1024-ensemble targets from seeds 100-139 (overrides --targets, --crop, --in-ens, "
                         "--tgt-ens, --seeds; trains 2000 steps of 8 frames)")
    a = ap.parse_args(argv)
    if a.config5:
        a.targets, a.crop, a.in_ens, a.tgt_ens, a.seeds = 2000, 0, "4", 1024, "100-139"
        a.steps, a.batch = 2000, 8
    lo, hi = (int(x) for x in a.seeds.split("-"))
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731
    d = Denoiser(seed=1)
    if a.init:
        d.load(a.init)
    t0 = time.perf_counter()
    dist = tuple(float(x) for x in a.dist.split(","))
    build_pairs(d, list(range(lo, hi + 1)), a.targets, a.crop, tuple(int(x) for x in a.in_ens.split(",")), a.tgt_ens,
                a.splats, rng_seed=a.rng_seed, log=log, dist_range=dist, time_budget=a.time_budget)
    t1 = time.perf_counter()
    losses = train(d, a.steps, a.batch, a.lr, log=log)
    t2 = time.perf_counter()
    d.save(a.out)
    log(f"[done] pairs {t1 - t0:.1f} s, training {t2 - t1:.1f} s, final loss {losses[-50:].mean():.5g} -> {a.out}")
    d.close()


if __name__ == "__main__":
    main()
