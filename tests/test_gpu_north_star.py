"""The headline workload against the oracle, in the configuration the benchmark times.

Workload: 560,000 clustered Gaussians (seed 1, SH degree 3), 1280x720, SPP 1, per-splat cap 16,384, orbit cameras at
radius 2 / 6 / 15 (near / mid / far); PBVR Extinction C3+R (the hot path, crates/gps-render/src/gps.rs:861-925 over
shaders/gps_splat.wgsl), GPS and screen occupancy on the same scene (the three methods the reference's bench times,
crates/gps-verify/src/main.rs:124-142).

The GPU runs with the launch tuning the autotuner picks for each view on an MI355X (the benchmark's tuning: 256 / 64
/ 16 workgroups per CU, 8 particles per thread, depth windows of 64 / 128 / 512 MiB, so up to 9 / 18 / 64 ensembles per
particle launch, the count-only path and the splat order on) and without the debug hook.  Against oracle/gps_mirror.c on the
host's threads: the accumulation buffer bit for bit and all nine counters.  A second pass keeps the per-ensemble
buffers (debug hook) for one ensemble and compares the per-splat counts, the exclusive offsets, the particle total and
the packed depth|id buffer.
"""
import ctypes
import os
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

gh = pytest.importorskip("gps_hip")
from gps_hip import scenes as S  # noqa: E402

W, H = 1280, 720
STAT_KEYS = ("points", "truncated_splats", "dropped_points", "visible_splats", "orphan_subpixels", "candidates",
             "undercovered_splats", "skipped_ensembles", "ensembles")
# the autotuner's choice per view on an MI355X for this workload (particle workgroups per CU, particles per thread,
# sort items per thread, ensembles per launch (0 = from the depth window), depth window MiB, count-only, block)
VIEW_TUNING = {
    "near": (256, 8, 64, 0, 64, 1, 256),
    "mid": (64, 8, 64, 0, 128, 1, 256),
    "far": (16, 8, 64, 0, 512, 1, 256),
}
_SCENE = {}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scene():
    if not _SCENE:
        _SCENE["g"], _SCENE["sh"] = S.clustered(560_000, seed=1, sh_degree=3)
        O.lib.mir_set_threads(int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1))
    return _SCENE["g"], _SCENE["sh"]


def tuning(view):
    v = VIEW_TUNING[view]
    return dict(particle_waves_per_cu=v[0], particles_per_thread=v[1], sort_items_per_thread=v[2],
                batch_ensembles=v[3], depth_window_mib=v[4], count_only=v[5], particle_block=v[6])


def mparams(p):
    m = O.MParams()
    ctypes.memmove(ctypes.byref(m), ctypes.byref(p), ctypes.sizeof(O.MParams))
    return m


def ocam(c):
    o = O.OrcCamera()
    ctypes.memmove(ctypes.byref(o), ctypes.byref(c), ctypes.sizeof(O.OrcCamera))
    return o


def method_params(method):
    if method == "ext_c3r":
        return gh.RenderParams(spp_side=1, method=gh.Method.Extinction, max_points_per_splat=16384,
                               pbvr=dict(calibration=3, radial_correction=1))
    if method == "gps":
        return gh.RenderParams(spp_side=1, method=gh.Method.Gps, max_points_per_splat=16384)
    return gh.RenderParams(spp_side=1, method=gh.Method.ScreenOccupancy, max_points_per_splat=16384)


def gpu_render(view, p, first_seed, count, keep=False):
    g, sh = scene()
    saved = gh.get_tuning()
    gh.set_tuning(**tuning(view))
    r = gh.GpsRenderer(W, H, 1)
    try:
        r.set_gaussians_sh(g, sh, 3)
        if keep:
            r.set_debug(gh.DEBUG_KEEP_BUFFERS)
        r.render_ensembles(S.clustered_camera(view, W, H), p, first_seed, count)
        out = dict(accum=r.read_accum(), stats=r.read_stats(), peaks=r.read_queue_peaks())
        if keep:
            out.update(r.read_counts(), depth_id=r.read_depth_id())
        return out
    finally:
        r.close()
        gh.set_tuning(**saved)


def stats_tuple(st):
    return tuple(int(getattr(st, k)) for k in STAT_KEYS)


def mirror(view, p, first_seed, count, keep=False):
    g, sh = scene()
    t0 = time.perf_counter()
    m = O.mirror_render(g, sh, 3, ocam(S.clustered_camera(view, W, H)), mparams(p), W, H, first_seed, count,
                        want_buffers=keep)
    say(f"    oracle: {time.perf_counter() - t0:.1f} s")
    return m


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("method", ["ext_c3r", "gps", "occupancy"])
@pytest.mark.parametrize("view", ["far", "mid", "near"])
def test_config2_bit_exact_in_timed_configuration(view, method):
    """Four ensembles in one call, as the benchmark issues them: accumulation buffer and all counters."""
    p = method_params(method)
    say(f"[config2 {view} {method}] GPU, 4 ensembles, tuning {VIEW_TUNING[view]}")
    gpu = gpu_render(view, p, 1, 4)
    m = mirror(view, p, 1, 4)
    st = gpu["stats"]
    say(f"    points {st.points} candidates {st.candidates} dropped {st.dropped_points} "
        f"truncated {st.truncated_splats} visible {st.visible_splats}")
    assert stats_tuple(st) == tuple(int(x) for x in m["stats"][:9])
    assert st.points > 0 and st.skipped_ensembles == 0 and st.ensembles == 4
    assert np.array_equal(gpu["accum"].view(np.uint32), m["accum"].view(np.uint32))
    # the count-only LDS buffers stayed within their capacities (per wave: under 64 waiting plus one round's 64)
    q, qcap, wk, wcap = gpu["peaks"]
    say(f"    count-only queue peak {q} / {qcap}, work buffer peak {wk} / {wcap}")
    assert q <= qcap and wk < wcap
    if method == "ext_c3r" and view != "far":
        assert wk > 0 and q > 0   # the path ran


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("method", ["ext_c3r", "gps"])
@pytest.mark.parametrize("view", ["far", "mid", "near"])
def test_config2_buffers_bit_exact(view, method):
    """One ensemble with the per-ensemble buffers kept: per-splat counts, exclusive offsets, particle total and the
    packed depth|id buffer (the reference's depth and winner buffers, gps_splat.wgsl:28-29)."""
    p = method_params(method)
    say(f"[config2 buffers {view} {method}] GPU, 1 ensemble, debug hook")
    gpu = gpu_render(view, p, 9, 1, keep=True)
    m = mirror(view, p, 9, 1, keep=True)
    assert stats_tuple(gpu["stats"]) == tuple(int(x) for x in m["stats"][:9])
    assert np.array_equal(gpu["counts"], m["counts"])
    assert np.array_equal(gpu["offsets"], m["offsets"])
    assert gpu["total"] == m["total"] and gpu["skipped"] == m["skipped"] == 0
    assert np.array_equal(gpu["depth_id"], m["depth_id"])
    assert np.array_equal(gpu["accum"].view(np.uint32), m["accum"].view(np.uint32))
