"""The viewer policies (gps_hip/viewer.py) driving the real renderer: a converged history equals one direct
render_ensembles call bit for bit, and the Auto path's demand probe and the sorted path run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gh = pytest.importorskip("gps_hip")
from gps_hip import viewer as V  # noqa: E402


def test_frames_accumulate_the_same_history_as_one_call():
    v = V.Viewer(160, 120)
    try:
        v.load_demo()
        v.configure(160, 120, 2, 1.0, per_frame=4, target=12, max_points=16384)
        v.set_render_path(0, 30.0)
        t = 0.0
        for _ in range(6):
            t += 16.7
            v.frame(t)
        assert v.accumulated() == 12 and v.image.shape == (120, 160, 4)
        acc = v.renderer.read_accum()
        r = gh.GpsRenderer(160, 120, 2)
        r.set_gaussians(V.demo_gaussians())
        r.render_ensembles(v.cam.to_camera(160, 120), v.params, 1, 12)
        assert np.array_equal(acc.view(np.uint32), r.read_accum().view(np.uint32))
        assert np.array_equal(v.image, r.composite(v.params, 160, 120))
        r.close()
        # the viewer's history and presented frame against the oracle
        import ctypes
        from oracle import oracle as O
        m = O.MParams()
        ctypes.memmove(ctypes.byref(m), ctypes.byref(v.params), ctypes.sizeof(O.MParams))
        oc = O.OrcCamera()
        ctypes.memmove(ctypes.byref(oc), ctypes.byref(v.cam.to_camera(160, 120)), ctypes.sizeof(O.OrcCamera))
        ref = O.mirror_render(V.demo_gaussians(), None, 0, oc, m, 160, 120, 1, 12)
        assert np.array_equal(acc.view(np.uint32), ref["accum"].view(np.uint32))
        assert np.array_equal(v.image, O.mirror_composite(ref["accum"], v.params.exposure, v.params.srgb_output,
                                                          160, 120))
    finally:
        v.renderer.close()


def test_auto_path_probes_and_the_sorted_path_renders():
    v = V.Viewer(640, 480)
    try:
        v.load_demo()
        v.configure(640, 480, 2, 1.0, per_frame=4, target=8, max_points=16384)
        v.set_render_path(2, 0.1)          # a threshold of 0.1M points: the demo's demand (2.5M) is above it
        # the probe's demand arrives with a readback issued after it (asynchronous, gps.rs:1005-1038): the GPU
        # idles between frames, as a frame's time would let it, and the first or second frame decides
        t = 16.7
        for _ in range(4):
            v.frame(t)
            t += 16.7
            v.renderer.wait_idle()
            if v.decided_epoch == v.history_epoch:
                break
        assert v.last_demand > 0 and v.sorted_active and v.accumulated() == 8
        v.set_render_path(2, 100000.0)     # far above it: GPS particles
        for _ in range(6):
            v.renderer.wait_idle()
            v.frame(t)
            t += 16.7
            if v.accumulated() == 8 and not v.sorted_active:
                break
        assert not v.sorted_active and v.accumulated() == 8
    finally:
        v.renderer.close()
