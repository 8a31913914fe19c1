"""Benchmark of the particle-rendering hot path on one MI355X.

Workload (one "step" = one render_ensembles call): the synthetic 560,000-Gaussian clustered scene (seed 1,
SH degree 3), 1280x720, SPP 1, PBVR Extinction C3+R, per-splat cap 16,384, 64 ensembles, orbit camera at
radius 2 / 6 / 15 (near / mid / far).  The headline `value` is GPU ms per ensemble at the chosen view
(default near); all three views are reported in `config`.  Inputs are resident in device memory before
the timed region.

The `roofline` object describes the kernel with the largest share of GPU time, from HIP events recorded
around every kernel launch on the renderer's stream during the timed steps.  The `cpu_baseline` object
times the CPU oracle's f32 restatement of the same pipeline (single thread) on a bounded sample.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gps-pbvr-web_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "GPU ms to PSNR≥40dB vs 1024-ens ref (batched+denoised vs plain); ms/ensemble"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, n, n_vis, sh_floats, particles, S, X, ens):
    """Algorithmic HBM bytes of one timed launch set (summed over the timed region)."""
    if name == "view":
        return n * (64 + 4 * sh_floats + 96 + 16 + 8)
    if name == "count_scan":
        return ens * n * (8 + 4)
    if name == "particles":
        return particles * 16 + ens * n_vis * 100
    if name == "resolve":
        return ens * S * (8 + 8 + 16) + 32 * X
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--view", default="near", choices=["near", "mid", "far"])
    ap.add_argument("--views", default="near,mid,far")
    ap.add_argument("--ensembles", type=int, default=64)
    ap.add_argument("--cpu-fraction", type=int, default=32, help="CPU baseline uses 1/F of the splats")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="bench", choices=["bench", "calib"],
                    help="calib: one exclusive scan of 64 Mi u32 (reads and writes exactly 256 MiB), the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import gps_hip as gh
    from gps_hip import scenes as S

    if a.mode == "calib":
        x = np.ones(64 << 20, np.uint32)
        y = gh.scan_exclusive_u32(x)
        assert int(y[-1]) == (64 << 20) - 1
        print(json.dumps({"calib_bytes_read": 4 * x.size, "calib_bytes_written": 4 * x.size}), flush=True)
        return

    W, H = 1280, 720
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    n = len(g)
    p = gh.RenderParams(spp_side=1, method=gh.Method.Extinction, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1))
    r = gh.GpsRenderer(W, H, 1)
    r.set_gaussians_sh(g, sh, 3)
    views = [v for v in a.views.split(",") if v]
    if a.view not in views:
        views.append(a.view)
    per_view = {}
    prof_view = {}
    for v in views:
        cam = S.clustered_camera(v, W, H)
        seed = 1
        for _ in range(a.warmup):
            r.reset_accum()
            r.render_ensembles(cam, p, seed, a.ensembles)
            seed += a.ensembles
        r.wait_idle()
        r.reset_stats()
        r.profile(True)
        r.wait_idle()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            r.reset_accum()
            r.render_ensembles(cam, p, seed, a.ensembles)
            seed += a.ensembles
        r.wait_idle()
        wall = time.perf_counter() - t0
        prof = r.profile_read()
        r.profile(False)
        st = r.read_stats()
        ens = a.steps * a.ensembles
        gpu_ms = sum(t for _, t in prof.values())
        particles = st.candidates - st.dropped_points
        per_view[v] = dict(ms_per_ensemble=gpu_ms / ens, wall_ms_per_ensemble=1000.0 * wall / ens,
                           points_per_ensemble=st.points / ens, particles_per_ensemble=particles / ens,
                           candidates_per_ensemble=st.candidates / ens, truncated_per_ensemble=st.truncated_splats / ens,
                           skipped=st.skipped_ensembles, visible_splats=st.visible_splats / ens)
        prof_view[v] = (prof, st, ens, particles)
        log(f"[{v}] {per_view[v]}")
        log(f"[{v}] kernels: " + ", ".join(f"{k}={t / c:.3f}ms x{c}" for k, (c, t) in prof.items()))

    head = a.view
    prof, st, ens, particles = prof_view[head]
    n_vis = st.visible_splats / max(ens, 1)
    Sx = W * H
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dcount, dms) = dom
    dbytes = kernel_bytes(dname, n, n_vis, 45, particles, Sx, Sx, ens)
    achieved = dbytes / (dms * 1e-3) / 1e9 if dms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "avg_launch_ms": round(dms / dcount, 4), "algorithmic_bytes_per_launch": int(dbytes / dcount)}
    kernels = {}
    for k, (c, t) in prof.items():
        b = kernel_bytes(k, n, n_vis, 45, particles, Sx, Sx, ens)
        kernels[k] = {"launches": c, "avg_ms": round(t / c, 4),
                      "GBps": round(b / (t * 1e-3) / 1e9, 1) if t > 0 and b else None}

    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        frac = max(1, a.cpu_fraction)
        m = max(1, n // frac)
        cam = S.clustered_camera(head, W, H)
        mp = O.MParams()
        ctypes.memmove(ctypes.byref(mp), ctypes.byref(p), ctypes.sizeof(mp))
        oc = O.OrcCamera()
        ctypes.memmove(ctypes.byref(oc), ctypes.byref(cam), ctypes.sizeof(oc))
        t0 = time.perf_counter()
        O.mirror_render(g[:m], sh[:m * 45], 3, oc, mp, W, H, 1, 1)
        dt = time.perf_counter() - t0
        model = ""
        try:
            model = subprocess.run(["bash", "-c", "lscpu | grep 'Model name' | head -1"], capture_output=True,
                                   text=True).stdout.split(":", 1)[-1].strip()
        except Exception:
            pass
        cpu = {"value": round(dt * 1000.0 * frac, 1), "unit": "ms/ensemble", "cores": 1, "kind": "port",
               "sample": f"f32 oracle of the same pipeline, first {m} of {n} splats (1/{frac}), 1 ensemble, "
                         f"{head} view, scaled x{frac}; host {model}, {os.cpu_count()} cpus visible",
               "sample_seconds": round(dt, 2)}
        log(f"[cpu] {cpu}")

    v = per_view[head]
    out = {
        "metric": METRIC,
        "value": round(v["ms_per_ensemble"], 4),
        "unit": "ms/ensemble",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(v["ms_per_ensemble"] * a.ensembles, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "clustered 560k Gaussians SH3, 1280x720, PBVR Extinction C3+R, SPP 1, cap 16384, "
                        f"{a.ensembles} ensembles per step, {head} orbit camera",
            "views": per_view,
            "kernels": kernels,
            "gaussians_per_s": round(n * a.ensembles / (v["ms_per_ensemble"] * a.ensembles * 1e-3), 1),
            "particles_per_s": round(v["particles_per_ensemble"] / (v["ms_per_ensemble"] * 1e-3), 1),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
