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
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gps-pbvr-web_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "GPU ms to PSNR≥40dB vs 1024-ens ref (batched+denoised vs plain); ms/ensemble"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, n, sh_floats, hits, S, X, ens, launches):
    """Algorithmic HBM bytes of all `launches` launches of kernel `name` in the timed region (`ens` ensembles,
    `hits` particles that reached a subpixel, over all of them)."""
    if name == "view":        # per render call: splat + SH read, 6 x float4 cache + colour + count info written
        return launches * n * (64 + 4 * sh_floats + 96 + 16 + 8)
    if name == "count_scan":  # per splat and ensemble: count info read, offset written
        return ens * n * (8 + 4)
    if name == "particles":   # 8-byte read-modify-write per particle that reaches a subpixel (the others write
        # nothing); per launch the splats' 96-byte view records and splat order, per ensemble their offsets
        return hits * 16 + launches * n * (96 + 4) + ens * n * 4
    if name == "resolve":     # depth|id read + reset, winner colour per subpixel; accumulation read + write
        return ens * S * (8 + 8 + 16) + launches * 32 * X
    return 0


# The autotuner's choice per view of this workload on an MI355X (bench.py --autotune): particle workgroups per CU,
# particles per thread, sort items per thread, ensembles per launch (0: from the depth window), depth window MiB,
# count-only path, workgroup size, sort digit bits.  The default run uses it, so that its launches are the ones
# the committed rocprofv3 profiles were taken on.
VIEW_TUNING = {
    "near": (256, 8, 64, 0, 64, 1, 256, 8),
    "mid": (64, 8, 64, 0, 128, 1, 256, 8),
    "far": (16, 8, 64, 0, 512, 1, 256, 8),
}


def set_view_tuning(gh, view):
    t = VIEW_TUNING[view]
    gh.set_tuning(particle_waves_per_cu=t[0], particles_per_thread=t[1], sort_items_per_thread=t[2],
                  batch_ensembles=t[3], depth_window_mib=t[4], count_only=t[5], particle_block=t[6],
                  sort_digit_bits=t[7])


def particle_roofline(dname, dcount, dms, n, ens, particles, hits, Sx, view):
    """The dominant kernel's roofline.  k_particles is bound by VALU issue: its rate is the lane-instructions per
    generated particle that rocprofv3 counted (SQ_INSTS_VALU x 64, profiles/particles_valu.json, on this workload
    and tuning) times this run's particles per launch, over this run's HIP-event launch time, against the fastest
    VALU rate gps_valu_bench.hip reached (counted the same way).  Its HBM side: bytes per launch measured by
    FETCH_SIZE / WRITE_SIZE (profiles/particles_traffic.json) and the algorithmic bytes (kernel_bytes), both over
    the same launch time."""
    avg_s = dms * 1e-3 / dcount
    alg = kernel_bytes(dname, n, 45, hits, Sx, Sx, ens, dcount) / dcount
    out = {"kernel": dname, "avg_launch_ms": round(dms / dcount, 4)}
    vpath = os.path.join(ROOT, "profiles", "particles_valu.json")
    tpath = os.path.join(ROOT, "profiles", "particles_traffic.json")
    if dname != "particles" or not os.path.exists(vpath):
        out.update(bound="hbm", achieved=round(alg / avg_s / 1e9, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=round(alg / avg_s / 1e9 / HBM_PEAK_GBS, 5), traffic=None)
        return out
    vp = json.load(open(vpath))
    lane_per_particle = vp["views"][view]["lane_insts_per_particle"]
    ppl = particles / dcount
    achieved = lane_per_particle * ppl / avg_s / 1e12
    peak = vp["ceiling"]["t_lane_insts_per_s"]
    out.update(bound="valu", achieved=round(achieved, 3), peak=peak, unit="T VALU lane-instructions/s",
               frac=round(achieved / peak, 4), traffic=None,
               lane_insts_per_particle=lane_per_particle, particles_per_launch=int(ppl),
               valu_port_busy=vp["views"][view].get("issue", {}).get("valu_port_busy"),
               floor_lane_insts_per_particle=vp["floor_lane_insts_per_particle"],
               source="profiles/particles_valu.json (SQ_INSTS_VALU per particle; ceiling from bench.py --mode valu)")
    hbm = {"algorithmic_bytes_per_launch": int(alg), "algorithmic_GBps": round(alg / avg_s / 1e9, 2),
           "peak": HBM_PEAK_GBS}
    if os.path.exists(tpath):
        tr = json.load(open(tpath))
        if view in tr.get("views", {}):
            t = tr["views"][view]["hbm_bytes_per_ensemble"] * ens / dcount
            out["traffic"] = int(t)
            hbm.update(measured_GBps=round(t / avg_s / 1e9, 2), frac=round(t / avg_s / 1e9 / HBM_PEAK_GBS, 5),
                       source="profiles/particles_traffic.json (FETCH_SIZE, WRITE_SIZE, calibrated)")
    out["hbm"] = hbm
    return out


def psnr(a, b):
    """crates/gps-core/src/oracle.rs:503-514: MSE over all channels, peak 1, capped at 120 dB."""
    d = (a.astype(np.float64) - b.astype(np.float64)) ** 2
    mse = float(d.mean())
    return 120.0 if mse <= 1e-20 else 10.0 * math.log10(1.0 / mse)


def reference_image(r, cam, p, ref_ensembles=1024, ref_seed=1_000_000):
    """The yardstick of the PSNR legs: an independent ref_ensembles-ensemble plain render of the view."""
    r.reset_accum()
    r.render_ensembles(cam, p, ref_seed, ref_ensembles)
    acc = r.read_accum()
    return acc[..., :3] / np.maximum(acc[..., 3:4], 1.0)


def time_to_psnr(r, cam, p, ms_per_ens, ref, target_db=40.0):
    """Plain ensemble averaging: smallest ensemble count whose mean reaches target_db against `ref`, and the
    GPU time that count costs."""
    from gps_hip.denoise import ladder
    r.reset_accum()
    done, curve = 0, []
    for e in ladder(1):
        r.render_ensembles(cam, p, 1 + done, e - done)
        done = e
        acc = r.read_accum()
        img = acc[..., :3] / np.maximum(acc[..., 3:4], 1.0)
        q = psnr(img, ref)
        curve.append((e, round(q, 2)))
        if q >= target_db:
            return {"ensembles": e, "gpu_ms": round(e * ms_per_ens, 3), "psnr_db": round(q, 2), "curve": curve}
    return {"ensembles": None, "gpu_ms": None, "psnr_db": curve[-1][1], "curve": curve}


def bench_atomics(a):
    """64-bit atomicMin throughput (gps_atomic_bench.hip) against which the particle kernel's L2 atomic rate is
    read: buffers of 8 MiB (one 1280x720 ensemble's depth|id buffer), 32 MiB (a four-ensemble launch), 256 MiB
    and 2 GiB; address windows per wave and iteration of 1 (all lanes on one address), 64, 1024, 16384 entries
    and the whole buffer; plain atomics and write_min's load-then-atomic form."""
    lib = ctypes.CDLL(os.path.join(ROOT, "gps-pbvr-web_amd", "libgps_atomicbench.so"))
    lib.gps_atomic_bench.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    blocks, per_thread, reps = 4096, 128, 5
    n_atomics = blocks * 256 * per_thread
    rows = []
    for log2_n in (20, 22, 25, 28):
        for log2_w in (0, 6, 10, 14, log2_n):
            for mode in (0, 1):
                ms = ctypes.c_float()
                rc = lib.gps_atomic_bench(log2_n, log2_w, mode, blocks, per_thread, reps, ctypes.byref(ms))
                if rc != 0:
                    raise RuntimeError(f"gps_atomic_bench({log2_n}, {log2_w}, {mode}) returned {rc}")
                row = dict(buffer_mib=(8 << log2_n) >> 20, window=1 << log2_w, mode=("atomic", "load+atomic")[mode],
                           ms=round(ms.value, 4), g_atomics_per_s=round(n_atomics / ms.value / 1e6, 2))
                log(f"[atomics] {row}")
                rows.append(row)
    print(json.dumps({"metric": "64-bit atomicMin throughput", "unit": "G atomics/s", "atomics_per_launch": n_atomics,
                      "launch": f"{blocks} x 256 threads x {per_thread} atomics", "reps": reps, "rows": rows}),
          flush=True)


def bench_valu(a):
    """VALU issue throughput (gps_valu_bench.hip): 8 independent chains per thread of f32 FMAs, of the particle
    streams' PCG hash, of its add/xor/shift part alone, and of single instruction classes, at full occupancy
    (65,536 workgroups of 256 threads).
    Run under rocprofv3 --pmc SQ_INSTS_VALU to count the lane-instructions of each launch; the kernel's rate against
    the fastest of these is the particle kernel's VALU roofline."""
    lib = ctypes.CDLL(os.path.join(ROOT, "gps-pbvr-web_amd", "libgps_atomicbench.so"))
    lib.gps_valu_bench.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_float)]
    blocks, iters, reps = 65536, 256, 5
    rows = []
    for mode, name in ((0, "fma_f32"), (1, "pcg_hash"), (2, "add_xor_shift"), (3, "add_u32"), (4, "xor_b32"),
                       (5, "mul_lo_u32"), (6, "add_f32"), (7, "fma_f32+add_u32"), (8, "lshrrev_const"),
                       (9, "lshrrev_vgpr"), (10, "bfe_u32"), (11, "alignbit_b32"), (12, "lshl_add_u32"),
                       (13, "xad_u32"), (14, "mad_u32_u24"), (15, "mul_hi_u32"), (16, "lshl_or_b32"),
                       (17, "xor_b32+lshrrev_const"), (18, "pcg_hash_x2_asm"), (19, "pcg_hash_x4_asm"),
                       (20, "pcg_hash_x2_asm_literals"), (21, "exp_f32"), (22, "rcp_f32"), (23, "sqrt_f32"),
                       (24, "sin_f32"), (25, "log_f32"), (26, "exp_f32+fmac_f32")):
        ms = ctypes.c_float()
        rc = lib.gps_valu_bench(mode, blocks, iters, reps, ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(f"gps_valu_bench({mode}) returned {rc}")
        rows.append(dict(mode=name, ms=round(ms.value, 4), chain_steps_per_launch=blocks * 256 * 8 * iters))
        log(f"[valu] {rows[-1]}")
    print(json.dumps({"metric": "VALU issue throughput", "launch": f"{blocks} x 256 threads x 8 chains x {iters} steps",
                      "reps": reps, "rows": rows}), flush=True)


def bench_alpha(a):
    """The sorted alpha path (render_sorted: project, 24-bit radix sort of the splat keys, tile binning and its
    sort, per-tile f32 blending) on the clustered 560k-Gaussian SH3 scene at 1280x720, at the three orbit
    views: host wall time per frame (it includes the path's host read of the tile-pair count), then, in as many
    frames again, the HIP-event time of its profiled stages."""
    import gps_hip as gh
    from gps_hip import scenes as S
    W, H = 1280, 720
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    p = gh.RenderParams(spp_side=1)
    r = gh.GpsRenderer(W, H, 1)
    r.set_gaussians_sh(g, sh, 3)
    frames = max(a.steps, 1) * 4
    out = {}
    for v in [x for x in a.views.split(",") if x]:
        cam = S.clustered_camera(v, W, H)
        for _ in range(max(a.warmup, 1)):
            r.render_sorted(cam, p)
        r.wait_idle()
        t0 = time.perf_counter()
        for _ in range(frames):
            r.render_sorted(cam, p)
        r.wait_idle()
        wall = (time.perf_counter() - t0) * 1000.0 / frames
        # the per-stage breakdown in frames of their own: its events add gaps between the stages
        r.profile(True)
        for _ in range(frames):
            r.render_sorted(cam, p)
        r.wait_idle()
        prof = r.profile_read()
        r.profile(False)
        out[v] = {"wall_ms_per_frame": round(wall, 3),
                  "profiled_gpu_ms_per_frame": round(sum(t for _, t in prof.values()) / frames, 3),
                  "kernels": {k: {"launches": c, "avg_ms": round(t / c, 4)} for k, (c, t) in prof.items()}}
        log(f"[alpha {v}] {out[v]}")
    print(json.dumps({"mode": "alpha", "workload": "clustered 560k Gaussians SH3 seed 1, 1280x720, sorted alpha path",
                      "frames_per_view": frames, "views": out}), flush=True)
    r.close()


def bench_sort3m(a):
    """3,000,000 clustered Gaussians (seed 2, SH3), 1920x1080, SPP side 2, GPS, cap 65,536: the packed-minimum
    resolve and the depth-sorted resolve (64-bit onesweep sort of every particle) per view, the sorted one at
    each radix digit width (4..8 bits; 8 is the default)."""
    import gps_hip as gh
    from gps_hip import scenes as S
    W, H = 1920, 1080
    g, sh = S.clustered(3_000_000, seed=2, sh_degree=3)
    r = gh.GpsRenderer(W, H, 2)
    r.set_gaussians_sh(g, sh, 3)
    out = {}
    for v in [x for x in a.views.split(",") if x]:
        cam = S.clustered_camera(v, W, H)
        res = {}
        saved = gh.get_tuning()
        for path, bits in ((1, 8), (2, 8), (2, 7), (2, 6), (2, 5), (2, 4)):
            gh.set_tuning(sort_digit_bits=bits)
            p = gh.RenderParams(spp_side=2, method=gh.Method.Gps, max_points_per_splat=65536, path=path)
            ens = max(1, min(a.ensembles, 8))
            r.reset_accum()
            r.render_ensembles(cam, p, 1, 1)
            r.wait_idle()
            r.reset_stats()
            r.profile(True)
            r.render_ensembles(cam, p, 100, ens)
            r.wait_idle()
            prof = r.profile_read()
            r.profile(False)
            st = r.read_stats()
            total = sum(t for _, t in prof.values())
            parts = (st.points - st.dropped_points) / ens
            d = {"ms_per_ensemble": round(total / ens, 3), "particles_per_ensemble": parts,
                 "skipped": st.skipped_ensembles,
                 "kernels": {k: {"launches": c, "avg_ms": round(t / c, 4)} for k, (c, t) in prof.items()}}
            if path == 2 and "sort_pass" in prof:
                c, t = prof["sort_pass"]
                keys = parts * ens
                d["sort_passes_per_ensemble"] = c / ens
                d["sort_pass_GBps"] = round(24.0 * keys * (c / ens) / (t * 1e-3) / 1e9, 1)
                d["sort_keys_per_s"] = round(keys / ((t + prof.get("sort_hist", (0, 0))[1]) * 1e-3), 1)
            res["atomic" if path == 1 else ("sorted" if bits == 8 else f"sorted_{bits}bit")] = d
            log(f"[sort3m {v} path={path} digit bits {bits}] {d}")
        gh.set_tuning(**saved)
        out[v] = res
    print(json.dumps({"mode": "sort3m", "workload": "clustered 3M Gaussians SH3 seed 2, 1920x1080, SPP side 2, GPS, "
                      "cap 65536", "views": out}), flush=True)
    r.close()


BATCH_SIZES = (20_000, 50_000, 100_000, 250_000, 560_000, 1_000_000, 1_500_000, 2_000_000)
BATCH_RES = ((640, 360), (1280, 720), (1920, 1080))


def bench_batch8(a):
    """Eight clustered scenes (seeds 10-17, SH3) of uneven size; scene i renders at BATCH_RES[i % 3]; each
    scene gets a 16-view 360-degree orbit (radius 6, elevation 15 degrees) with 32 ensembles per view of PBVR
    Extinction C3+R (SPP 1, cap 16384).  Legs: the fused batch (every view in the same few launches per round of
    ensembles) recorded into one hipGraph and launched straight onto the stream; the per-item batch (each view's
    own launches, on up to four streams) the same two ways; one submission per ensemble (one render_ensembles
    call per ensemble, the reference's trajectory-export pattern).  Every leg's images are compared bit for bit."""
    import gps_hip as gh
    from gps_hip import scenes as S
    views, ens = 16, 32
    p = gh.RenderParams(spp_side=1, method=gh.Method.Extinction, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1))
    rends, items, total_g = [], [], 0
    for i, n in enumerate(BATCH_SIZES):
        w, h = BATCH_RES[i % 3]
        g, sh = S.clustered(n, seed=10 + i, sh_degree=3)
        r = gh.GpsRenderer(w, h, 1)
        r.set_gaussians_sh(g, sh, 3)
        rends.append(r)
        for v in range(views):
            cam = gh.orbit_camera((0.0, 0.0, 0.0), 6.0, 2.0 * math.pi * v / views, math.radians(15.0),
                                  (0.0, 0.0, 1.0), S.FOV_60, w, h)
            items.append((r, cam, p, 1 + 1000 * v, ens))
        total_g += n * views * ens
        log(f"[batch8] scene {i}: {n} Gaussians at {w}x{h}")
    tuning = gh.get_tuning()
    if not a.no_autotune:
        tuning, tms = rends[4].autotune(items[4 * views][1], p, trial_ensembles=4)
        log(f"[batch8 autotune] {tuning} {tms:.3f} ms/ens")
    b = gh.Batch()
    out = {}
    last = [len(items) - 1 - k * views for k in range(len(BATCH_SIZES))][::-1]
    snap = None
    for leg, use_graph, per_item in (("graph", True, False), ("stream", False, False),
                                     ("per_item_graph", True, True), ("per_item_stream", False, True)):
        for _ in range(max(a.warmup, 1)):
            b.render(items, use_graph=use_graph, per_item=per_item)
        gms, rec = [], []
        for _ in range(max(a.steps, 1)):
            b.render(items, use_graph=use_graph, per_item=per_item)
            t = b.timing()
            gms.append(t["gpu_ms"])
            rec.append(t["record_ms"])
        out[leg] = {"gpu_ms": round(min(gms), 2), "record_ms": round(min(rec), 2), "graph_nodes": t["graph_nodes"]}
        imgs = {i: b.read(i) for i in last}
        if snap is None:
            snap = imgs
        else:
            out[leg]["bit_identical_to_graph"] = all(np.array_equal(snap[i].view(np.uint32), imgs[i].view(np.uint32))
                                                     for i in last)
        log(f"[batch8 {leg}] {out[leg]}")
    for r in rends:
        r.wait_idle()
    t0 = time.perf_counter()
    for r, cam, q, seed, count in items:
        r.reset_accum()
        for e in range(count):
            r.render_ensembles(cam, q, seed + e, 1)
    for r in rends:
        r.wait_idle()
    per_ens_ms = 1000.0 * (time.perf_counter() - t0)
    exact = all(np.array_equal(snap[i].view(np.uint32), items[i][0].read_accum().view(np.uint32)) for i in last)
    out["per_ensemble_submits"] = {"wall_ms": round(per_ens_ms, 2), "bit_identical_to_graph": bool(exact)}
    log(f"[batch8 per-ensemble] {out['per_ensemble_submits']}")
    # the same graph with the conservative cull (bit-identical images)
    for r in rends:
        r.set_cull(True)
    b.render(items, use_graph=True)
    t = b.timing()
    same = all(np.array_equal(snap[i].view(np.uint32), b.read(i).view(np.uint32)) for i in last)
    out["graph_cull"] = {"gpu_ms": round(t["gpu_ms"], 2), "bit_identical_to_graph": bool(same)}
    log(f"[batch8 graph+cull] {out['graph_cull']}")
    for r in rends:
        r.set_cull(False)
    gm = out["graph"]["gpu_ms"]
    print(json.dumps({
        "mode": "batch8",
        "workload": "8 clustered scenes (seeds 10-17; 20k..2M Gaussians, SH3) x 16-view orbit (r 6, elev 15 deg), "
                    "640x360/1280x720/1920x1080 cycling per scene, PBVR Extinction C3+R, SPP 1, cap 16384, "
                    "32 ensembles per view",
        "views": len(items), "ensembles": len(items) * ens, "legs": out,
        "ms_per_view": round(gm / len(items), 3), "ensembles_per_s": round(len(items) * ens / (gm * 1e-3), 1),
        "gaussian_ensembles_per_s": round(total_g / (gm * 1e-3), 1),
        "graph_speedup_vs_per_ensemble_submits": round(per_ens_ms / gm, 3),
        "fused_speedup_vs_per_item_graph": round(out["per_item_graph"]["gpu_ms"] / gm, 3), "tuning": tuning}),
        flush=True)
    b.close()
    for r in rends:
        r.close()


def bench_ply(a):
    """gps-verify bench-ply (crates/gps-verify/src/main.rs:104-143): per-ensemble GPU time on a 3DGS PLY file for
    GPS, PBVR Extinction C3+R and screen occupancy at three framed views (the robust sphere at x1, x1.85 and
    x3.2 distance, up -Y, yaw 0.6, pitch 0.25, fov 0.9), 6 single-ensemble renders each after a warm-up."""
    import gps_hip as gh
    path = a.ply or os.path.join(ROOT, "tests", "golden", "synthetic_sh1.ply")
    w, h = (int(x) for x in a.size.split("x"))
    g = gh.load_ply(path)
    gpu = g.to_gpu()
    r = gh.GpsRenderer(w, h, a.spp)
    if g.sh_degree > 0:
        r.set_gaussians_sh(gpu, g.sh_rest, g.sh_degree)
    else:
        r.set_gaussians(gpu)
    c, rad = gh.robust_sphere(gpu["pos_opacity"][:, :3])
    log(f"[bench-ply] {path}: {g.count} Gaussians, SH degree {g.sh_degree}, {w}x{h} spp side {a.spp}")
    env_max = int(os.environ.get("MAXPTS", "16384"))
    rows = []
    for vname, mul in (("near (x1.25)", 1.0), ("mid (x2.3)", 1.85), ("far (x4)", 3.2)):
        cam = gh.orbit_camera(tuple(c), gh.frame_sphere_distance(rad) * mul, 0.6, 0.25, (0.0, -1.0, 0.0), 0.9, w, h)
        for mname, method, cal, radial in (("GPS", gh.Method.Gps, 3, 0), ("PBVR Ext C3+R", gh.Method.Extinction, 3, 1),
                                           ("Screen occupancy", gh.Method.ScreenOccupancy, 0, 0)):
            p = gh.RenderParams(spp_side=a.spp, max_points_per_splat=env_max, method=method,
                                pbvr=dict(calibration=cal, radial_correction=radial))
            r.render_ensembles(cam, p, 1, 1)
            r.wait_idle()
            r.reset_stats()
            r.profile(True)
            reps = 6
            t0 = time.perf_counter()
            for i in range(reps):
                r.render_ensembles(cam, p, 100 + i, 1)
                r.wait_idle()
            wall = (time.perf_counter() - t0) * 1000.0 / reps
            prof = r.profile_read()
            r.profile(False)
            st = r.read_stats()
            row = {"view": vname, "method": mname, "gpu_ms_per_ensemble": round(sum(t for _, t in prof.values()) / reps, 4),
                   "wall_ms_per_ensemble": round(wall, 4), "points_per_ensemble": st.points // reps,
                   "truncated_per_ensemble": st.truncated_splats // reps}
            if method == gh.Method.ScreenOccupancy:
                row.update(candidates_per_ensemble=st.candidates // reps, skipped=st.skipped_ensembles,
                           orphan=st.orphan_subpixels)
            rows.append(row)
            log(f"[bench-ply] {row}")
    print(json.dumps({"mode": "bench-ply", "file": os.path.basename(path), "gaussians": g.count,
                      "sh_degree": g.sh_degree, "size": a.size, "spp_side": a.spp, "rows": rows}), flush=True)
    r.close()


# FLOPs of one denoiser forward per pixel: 3x3 convolutions 7->32, 3 x 32->32, 32->3 (multiply-adds x 2)
DENOISE_FLOP_PER_PIXEL = 2 * 9 * (7 * 32 + 3 * 32 * 32 + 32 * 3)
BF16_MFMA_PEAK_TFLOPS = 2500.0


def bench_denoise(a):
    """The learned denoiser on the 560k clustered scene's near view at 1280x720 (320 ensembles, features on):
    GPU ms per forward (features + 5 convolutions + fallback) from HIP events, layer by layer and fused, and the
    fused forward's MFMA roofline."""
    import gps_hip as gh
    from gps_hip import scenes as S
    from gps_hip import denoise as D
    W, H = 1280, 720
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    r = gh.GpsRenderer(W, H, 1)
    r.set_gaussians_sh(g, sh, 3)
    r.set_cull(True)
    r.set_features(True)
    r.render_ensembles(S.clustered_camera("near", W, H), D.params_ext_c3r(), 1, 320)
    d = gh.Denoiser()
    d.load(D.WEIGHTS)
    flop = DENOISE_FLOP_PER_PIXEL * W * H
    out = {"mode": "denoise", "workload": "denoiser forward, 1280x720, 7 feature channels, 5 layers",
           "flop_per_forward": flop}
    # the layer-by-layer kernels first, then the fused kernel apply runs by default (the same bits)
    for name, fused in (("layered", False), ("fused", True)):
        d.set_fused(fused)
        for _ in range(max(a.warmup, 1)):
            d.apply(r, fallback_ratio=4.0)
        ms = [d.apply(r, fallback_ratio=4.0)["gpu_ms"] for _ in range(max(a.steps, 1) * 10)]
        avg = float(np.mean(ms))
        tf = flop / (avg * 1e-3) / 1e12
        out[name] = {"forwards": len(ms), "avg_ms": round(avg, 4), "TFLOPs": round(tf, 2)}
    tf = out["fused"]["TFLOPs"]
    out["avg_ms"] = out["fused"]["avg_ms"]
    out["roofline"] = {"bound": "mfma", "achieved": tf, "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(tf / BF16_MFMA_PEAK_TFLOPS, 4)}
    print(json.dumps(out), flush=True)
    d.close()
    r.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--view", default="near", choices=["near", "mid", "far"])
    ap.add_argument("--views", default="near,mid,far")
    ap.add_argument("--ensembles", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-psnr", action="store_true", help="skip the time-to-40-dB measurement")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner per view instead of the committed per-view tuning (VIEW_TUNING)")
    ap.add_argument("--no-autotune", action="store_true", help="batch8: keep the default launch tuning")
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoised leg and the held-out evaluation")
    ap.add_argument("--no-cull", action="store_true", help="skip the timing of the culled renders")
    ap.add_argument("--heldout", type=int, default=10, help="held-out scenes (seeds 200..) for the denoiser")
    ap.add_argument("--tuning", default=None,
                    help="waves,ppt,items,batch,window[,count_only[,block[,digit_bits]]]: one fixed launch tuning "
                         "for every view (implies --no-autotune)")
    ap.add_argument("--ply", default=None, help="bench-ply: the 3DGS PLY file (default: the test fixture)")
    ap.add_argument("--size", default="480x270", help="bench-ply: WxH")
    ap.add_argument("--spp", type=int, default=1, help="bench-ply: SPP side")
    ap.add_argument("--mode", default="bench",
                    choices=["bench", "calib", "atomics", "valu", "alpha", "sort3m", "batch8", "denoise", "bench-ply"],
                    help="calib: one exclusive scan of 64 Mi u32 (reads and writes exactly 256 MiB), the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters; atomics: the 64-bit "
                         "atomicMin throughput microbenchmark")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); it has no multi-GPU mode",
              file=sys.stderr)
        sys.exit(2)

    import gps_hip as gh
    from gps_hip import scenes as S

    if a.tuning:
        v = [int(x) for x in a.tuning.split(",")]
        gh.set_tuning(particle_waves_per_cu=v[0], particles_per_thread=v[1], sort_items_per_thread=v[2],
                      batch_ensembles=v[3], depth_window_mib=v[4],
                      count_only=v[5] if len(v) > 5 else 1, particle_block=v[6] if len(v) > 6 else 256,
                      sort_digit_bits=v[7] if len(v) > 7 else 8)
        a.no_autotune = True
        a.autotune = False
    if a.mode == "calib":
        x = np.ones(64 << 20, np.uint32)
        y = gh.scan_exclusive_u32(x)
        assert int(y[-1]) == (64 << 20) - 1
        print(json.dumps({"calib_bytes_read": 4 * x.size, "calib_bytes_written": 4 * x.size}), flush=True)
        return

    if a.mode == "atomics":
        return bench_atomics(a)
    if a.mode == "valu":
        return bench_valu(a)
    if a.mode == "alpha":
        return bench_alpha(a)
    if a.mode == "sort3m":
        return bench_sort3m(a)
    if a.mode == "batch8":
        return bench_batch8(a)
    if a.mode == "denoise":
        return bench_denoise(a)
    if a.mode == "bench-ply":
        return bench_ply(a)

    W, H = 1280, 720
    g, sh = S.clustered(560_000, seed=1, sh_degree=3)
    n = len(g)
    p = gh.RenderParams(spp_side=1, method=gh.Method.Extinction, max_points_per_splat=16384,
                        pbvr=dict(calibration=3, radial_correction=1))
    p_gps = gh.RenderParams(spp_side=1, method=gh.Method.Gps, max_points_per_splat=16384)
    p_occ = gh.RenderParams(spp_side=1, method=gh.Method.ScreenOccupancy, max_points_per_splat=16384)
    r = gh.GpsRenderer(W, H, 1)
    r.set_gaussians_sh(g, sh, 3)
    views = [v for v in a.views.split(",") if v]
    if a.view not in views:
        views.append(a.view)
    tuning = {}
    per_view = {}
    prof_view = {}
    for v in views:
        cam = S.clustered_camera(v, W, H)
        if not a.autotune and not a.tuning:
            set_view_tuning(gh, v)
        elif a.autotune:
            # trials of about 4 ms of GPU work each (4 to 64 ensembles): light views get enough work per trial
            # for launch timing noise not to pick their tuning
            r.reset_accum()
            r.profile(True)
            r.render_ensembles(cam, p, 1, 4)
            r.wait_idle()
            est = sum(t for _, t in r.profile_read().values()) / 4.0
            r.profile(False)
            trial = int(min(64, max(4, math.ceil(4.0 / max(est, 1e-3)))))
            t, tms = r.autotune(cam, p, trial_ensembles=trial)
            log(f"[autotune {v}] {trial} ensembles per trial: {t} {tms:.3f} ms/ens")
        tuning[v] = gh.get_tuning()
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
        hits = r.read_particle_counters()["hits"]
        ens = a.steps * a.ensembles
        gpu_ms = sum(t for _, t in prof.values())
        particles = st.candidates - st.dropped_points
        per_view[v] = dict(ms_per_ensemble=gpu_ms / ens, wall_ms_per_ensemble=1000.0 * wall / ens,
                           points_per_ensemble=st.points / ens, particles_per_ensemble=particles / ens,
                           onscreen_particles_per_ensemble=hits / ens,
                           candidates_per_ensemble=st.candidates / ens, truncated_per_ensemble=st.truncated_splats / ens,
                           skipped=st.skipped_ensembles, visible_splats=st.visible_splats / ens)
        prof_view[v] = (prof, st, ens, particles, hits)
        # the same renders with the conservative cull: bit-identical images, fewer generated particles
        if a.no_cull:
            log(f"[{v}] {per_view[v]}")
            log(f"[{v}] kernels: " + ", ".join(f"{k}={t / c:.3f}ms x{c}" for k, (c, t) in prof.items()))
            continue
        r.set_cull(True)
        r.reset_accum()
        r.render_ensembles(cam, p, seed, a.ensembles)
        r.wait_idle()
        r.profile(True)
        for _ in range(a.steps):
            r.reset_accum()
            r.render_ensembles(cam, p, seed, a.ensembles)
            seed += a.ensembles
        r.wait_idle()
        cprof = r.profile_read()
        r.profile(False)
        r.set_cull(False)
        per_view[v]["ms_per_ensemble_cull"] = sum(t for _, t in cprof.values()) / ens
        # the reference's bench times GPS and screen occupancy on the same view (crates/gps-verify/src/main.rs:124-142)
        for mname, q in (("gps", p_gps), ("occupancy", p_occ)):
            r.reset_accum()
            r.render_ensembles(cam, q, seed, a.ensembles)
            r.wait_idle()
            r.reset_stats()
            r.profile(True)
            for _ in range(a.steps):
                r.reset_accum()
                r.render_ensembles(cam, q, seed, a.ensembles)
                seed += a.ensembles
            r.wait_idle()
            mprof = r.profile_read()
            r.profile(False)
            mst = r.read_stats()
            per_view[v][f"{mname}_ms_per_ensemble"] = sum(t for _, t in mprof.values()) / ens
            per_view[v][f"{mname}_points_per_ensemble"] = mst.points / ens
        log(f"[{v}] {per_view[v]}")
        log(f"[{v}] kernels: " + ", ".join(f"{k}={t / c:.3f}ms x{c}" for k, (c, t) in prof.items()))

    head = a.view
    prof, st, ens, particles, hits = prof_view[head]
    Sx = W * H
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dcount, dms) = dom
    roofline = particle_roofline(dname, dcount, dms, n, ens, particles, hits, Sx, head)
    kernels = {}
    for k, (c, t) in prof.items():
        b = kernel_bytes(k, n, 45, hits, Sx, Sx, ens, c)
        kernels[k] = {"launches": c, "avg_ms": round(t / c, 4),
                      "GBps": round(b / (t * 1e-3) / 1e9, 1) if t > 0 and b else None}

    # 64-bit atomic requests per ensemble counted by rocprofv3 (TCC_ATOMIC) on this workload, against the
    # atomicMin ceiling that gps_atomic_bench.hip measured
    apath = os.path.join(ROOT, "profiles", "particles_atomics.json")
    if head == "near" and "particles" in kernels and os.path.exists(apath):
        at = json.load(open(apath))
        kp = kernels["particles"]
        req = at["l2_atomic_requests_per_ensemble"] * ens / kp["launches"]
        per_s = req / (kp["avg_ms"] * 1e-3)
        ceil = at["ceiling"]["g_requests_per_s"] * 1e9
        kp["u64_atomics"] = {"requests_per_launch": int(req), "per_s": round(per_s), "ceiling_per_s": round(ceil),
                             "frac": round(per_s / ceil, 5), "source": "profiles/particles_atomics.json"}

    ttp, heldout = None, None
    if not a.no_psnr:
        from gps_hip import denoise as D
        den = None
        if not a.no_denoise:
            if not os.path.exists(D.WEIGHTS):
                raise SystemExit(f"denoiser weights missing: {D.WEIGHTS} (python -m gps_hip.denoise)")
            den = gh.Denoiser()
            den.load(D.WEIGHTS)
        # the PSNR legs render with the cull (the same images) and price each ensemble at the exact path's
        # measured GPU time
        r.set_cull(True)
        r.set_features(True)
        ttp = {}
        for v in views:
            cam = S.clustered_camera(v, W, H)
            ref = reference_image(r, cam, p)
            ms = per_view[v]["ms_per_ensemble"]
            leg = {"plain": time_to_psnr(r, cam, p, ms, ref), "denoised": None}
            if den is not None:
                leg["denoised"] = D.denoised_time_to_psnr(r, den, cam, p, ms, ref, fallback_ratio=4.0)
                if leg["plain"]["gpu_ms"] and leg["denoised"]["gpu_ms"]:
                    leg["speedup"] = round(leg["plain"]["gpu_ms"] / leg["denoised"]["gpu_ms"], 2)
            # the same ensemble counts priced at the culled renders' GPU time (bit-identical images; only the
            # generated-points counters differ), for callers that need the image alone
            cms = per_view[v].get("ms_per_ensemble_cull")
            if cms is not None:
                if leg["plain"]["ensembles"]:
                    leg["plain"]["gpu_ms_cull"] = round(leg["plain"]["ensembles"] * cms, 3)
                dn = leg["denoised"]
                if dn is not None and dn["ensembles"]:
                    dn["gpu_ms_cull"] = round(dn["ensembles"] * cms + dn["denoise_ms"], 3)
            ttp[v] = leg
            log(f"[psnr {v}] {leg}")
        r.set_features(False)
        r.set_cull(False)
        if den is not None and a.heldout > 0:
            rows = D.evaluate_heldout(den, seeds=range(200, 200 + a.heldout), log=log)
            keys = ("psnr_plain", "psnr_denoised", "max_abs_plain", "max_abs_denoised", "fallback_rate", "denoise_ms")
            heldout = {"scenes": len(rows), "input_ensembles": 4, "ref_ensembles": 1024,
                       **{f"mean_{k}": round(float(np.mean([x[k] for x in rows])), 4) for k in keys}, "rows": rows}
            den.close()

    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        mp = O.MParams()
        ctypes.memmove(ctypes.byref(mp), ctypes.byref(p), ctypes.sizeof(mp))
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)

        def ocam(view):
            oc = O.OrcCamera()
            ctypes.memmove(ctypes.byref(oc), ctypes.byref(S.clustered_camera(view, W, H)), ctypes.sizeof(oc))
            return oc

        def mirror_ms(view, nthreads):
            """the f32 oracle of the same pipeline on the whole scene, one ensemble (seed 1)"""
            O.lib.mir_set_threads(nthreads)
            t0 = time.perf_counter()
            m = O.mirror_render(g, sh, 3, ocam(view), mp, W, H, 1, 1)
            return 1000.0 * (time.perf_counter() - t0), m["stats"]

        head_ms, head_st = mirror_ms(head, threads)
        views_ms = {head: round(head_ms, 1)}
        for v in views:
            if v != head:
                views_ms[v] = round(mirror_ms(v, threads)[0], 1)
        far1_ms, _ = mirror_ms("far", 1)
        # the f64 port of the reference's own CPU particle renderer (oracle.rs:426-494), single-threaded like it:
        # PBVR Extinction C3 with the radial correction and per-particle seeds, base colour (it has no SH) and no
        # per-splat cap (it has none), on the whole scene at the far view, where no splat reaches the cap
        o64 = np.zeros((n, 14), np.float64)
        o64[:, 0:3] = g["pos_opacity"][:, :3]
        o64[:, 3:6] = g["log_scale"][:, :3]
        o64[:, 6:9] = g["rot_wxyz"][:, 1:4]
        o64[:, 9] = g["rot_wxyz"][:, 0]
        o64[:, 10] = g["pos_opacity"][:, 3]
        o64[:, 11:14] = g["color"][:, :3]
        popt = O.default_particle_options(level=3, rule=1, radial_correction=1, per_particle_seed=1)
        t0 = time.perf_counter()
        _, st64 = O.render_particles_3d(o64, ocam("far"), W, H, 1, 1, tuple(float(t) for t in p.background), 1, popt)
        f64_ms = 1000.0 * (time.perf_counter() - t0)
        model = ""
        try:
            model = subprocess.run(["bash", "-c", "lscpu | grep 'Model name' | head -1"], capture_output=True,
                                   text=True).stdout.split(":", 1)[-1].strip()
        except Exception:
            pass
        cpu = {"value": round(head_ms, 1), "unit": "ms/ensemble", "cores": threads, "kind": "port",
               "sample": f"f32 oracle of the same pipeline (particle loop over splats on {threads} OpenMP threads) on "
                         f"the whole {n}-splat scene, 1 ensemble (seed 1), {head} view, measured, not extrapolated: "
                         f"{int(head_st[5])} candidates, {int(head_st[0])} points; host {model}, {os.cpu_count()} cpus "
                         f"visible",
               "views_all_threads_ms": views_ms,
               "single_thread": {"value": round(far1_ms, 1), "cores": 1, "view": "far",
                                 "sample": "the same oracle on one thread, whole scene, 1 ensemble, far view"},
               "f64_reference_port": {"value": round(f64_ms, 1), "cores": 1, "view": "far",
                                      "sample": "f64 port of the reference's CPU particle renderer (render_particles_3d, "
                                                "per-particle seeds, base colour, no per-splat cap), whole scene, 1 "
                                                "ensemble, far view, where no splat reaches the GPU path's cap",
                                      "candidates": int(st64.candidates), "points": int(st64.points)}}
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
            "time_to_40db": ttp,
            "denoiser_heldout": heldout,
            "tuning": tuning,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
