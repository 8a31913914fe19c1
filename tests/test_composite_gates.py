"""The reference's composite gates (crates/gps-render/src/verify.rs:225-291) applied to the oracle's composite
(oracle/gps_mirror.c mir_composite), against the f64 expectation of tests/composite_gate.py.  The GPU composite is
held to the same gates in tests/test_gpu_gates.py and equals mir_composite bit for bit in tests/test_gpu_parity.py.

Planted mistakes (numpy variants of the composite, on the CPU) must fail the gates: the OETF before filtering, no
OETF, a flipped image, a missing half-pixel offset, and truncation instead of rounding."""
import numpy as np
import pytest

from composite_gate import REFERENCE_BOUND, STRICT, expected, gate_accum, probe_error, srgb_oetf, worst_error
from oracle import oracle as O
from scenes import BG, phase_scenes, to_gpu


def mirror_image(w, h, sets=16):
    """verify.rs:218-223: scene 1 ("single o=0.5"), 16 ensembles of RenderParams::default() with the phase-2
    background, spp side 2."""
    p = O.MParams(spp_side=2, density_scale=1.0, max_points_per_splat=1 << 16, near=0.05, exposure=1.0, srgb_output=1,
                  sh_degree=3, path=1, method=0)
    p.background[:] = BG
    p.pbvr.calibration = 3
    p.pbvr.low_pass_jitter = 1
    p.pbvr.base_k = 512.0
    p.pbvr.reference_pixel_length = 0.01
    return O.mirror_render(to_gpu(phase_scenes()[1][1]), None, 0, O.default_camera(200.0), p, w, h, 1, sets)["accum"]


def test_oetf_closed_forms():
    assert srgb_oetf(0.0) == 0.0 and srgb_oetf(1.0) == pytest.approx(1.0, abs=1e-15)
    assert srgb_oetf(0.0031308) == pytest.approx(12.92 * 0.0031308)
    # the two branches meet at the knee (continuity of the sRGB curve to 1e-7)
    assert abs(1.055 * 0.0031308 ** (1 / 2.4) - 0.055 - 12.92 * 0.0031308) < 1e-7
    assert srgb_oetf(0.5) == pytest.approx(0.7353569830524495, rel=1e-12)


def test_reference_gates_on_mirror_composite():
    """verify.rs:233-252 (five probes, same size) and 254-291 (2x upsample of the whole image)."""
    acc = mirror_image(96, 96)
    assert (acc[..., 3] == 16.0).all()
    same = O.mirror_composite(acc, 1.0, 1, 96, 96)
    assert probe_error(same, acc) <= REFERENCE_BOUND
    assert worst_error(same, acc) <= STRICT
    up = O.mirror_composite(acc, 1.0, 1, 192, 192)
    assert worst_error(up, acc) <= REFERENCE_BOUND
    assert worst_error(up, acc) <= STRICT
    assert (same[..., 3] == 255).all() and (up[..., 3] == 255).all()


@pytest.mark.parametrize("w,h,ow,oh,expo,gamma", [(96, 96, 96, 96, 1.0, 1), (96, 96, 192, 192, 1.0, 1),
                                                  (64, 40, 50, 70, 0.8, 1), (64, 40, 64, 40, 1.7, 0),
                                                  (33, 17, 100, 51, 1.3, 1), (50, 30, 25, 15, 1.0, 1)])
def test_mirror_composite_within_strict_bound(w, h, ow, oh, expo, gamma):
    """Any output size, exposure and gamma setting, on an accumulation that reaches every branch."""
    acc = gate_accum(w, h, seed=w * 7 + ow)
    img = O.mirror_composite(acc, expo, gamma, ow, oh)
    assert worst_error(img, acc, expo, bool(gamma)) <= STRICT


def _to8(e):
    return np.clip(np.rint(e), 0, 255).astype(np.uint8)


def _rgba(rgb):
    out = np.full(rgb.shape[:2] + (4,), 255, np.uint8)
    out[..., :3] = rgb
    return out


def test_planted_mistakes_fail_the_gates():
    acc = gate_accum(96, 96, seed=5)
    lin = acc[..., :3] / np.maximum(acc[..., 3:4], 1.0)
    # the correct composite written in numpy passes
    good = _rgba(_to8(expected(acc, 192, 192)))
    assert worst_error(good, acc) <= STRICT
    # OETF applied per texel before the bilinear filter (the double-gamma / wrong-order mistake)
    pre = acc.copy()
    pre[..., :3] = srgb_oetf(np.clip(lin, 0, 1)) * np.maximum(acc[..., 3:4], 1.0)
    wrong = _rgba(_to8(expected(pre, 192, 192, gamma=False)))
    assert worst_error(wrong, acc) > REFERENCE_BOUND
    # no OETF
    assert worst_error(_rgba(_to8(expected(acc, 192, 192, gamma=False))), acc) > REFERENCE_BOUND
    # vertically flipped
    assert worst_error(good[::-1].copy(), acc) > REFERENCE_BOUND
    # sampling at q = uv * size (half-pixel offset missing): shift the expectation by half an input pixel
    shifted = acc.copy()
    shifted[:, :-1] = acc[:, 1:]
    assert worst_error(_rgba(_to8(0.5 * (expected(acc, 192, 192) + expected(shifted, 192, 192)))), acc) > REFERENCE_BOUND
    # truncation instead of rounding to nearest passes the reference's 1.01 but not the strict bound
    trunc = _rgba(np.clip(np.floor(expected(acc, 96, 96)), 0, 255).astype(np.uint8))
    assert worst_error(trunc, acc) <= REFERENCE_BOUND
    assert worst_error(trunc, acc) > STRICT
