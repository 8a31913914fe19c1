"""f64 expectations of the display composite (shaders/gps_composite.wgsl:29-55), as the reference's phase-2 gate
states them (crates/gps-render/src/verify.rs:225-291): the linear pixel is sum / max(count, 1) (an f32 division,
verify.rs:245, 275), bilinear filtering at pixel centres with edge clamping happens in linear space, then exposure,
clamp to [0, 1], and one sRGB OETF (verify.rs:81-87), times 255.

The reference's bound on |8-bit value - expectation| is 1.01.  An RGBA8 target rounds to nearest, so a correct
composite also stays within 0.5 plus the f32 evaluation error (about 3e-4 of an 8-bit step here); STRICT states that
tighter bound, which a truncating conversion fails.
"""
import numpy as np

REFERENCE_BOUND = 1.01   # verify.rs:250, 289
STRICT = 0.51


def srgb_oetf(c):
    """verify.rs:81-87, f64"""
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(np.maximum(c, 0.0), 1.0 / 2.4) - 0.055)


def linear_image(accum):
    """(a[c] / max(a[3], 1)) as f64, with the division in f32 (verify.rs:245)."""
    a = np.ascontiguousarray(accum, np.float32)
    return (a[..., :3] / np.maximum(a[..., 3:4], np.float32(1.0))).astype(np.float64)


def expected(accum, out_w, out_h, exposure=1.0, gamma=True):
    """8-bit expectation for an out_w x out_h target (verify.rs:263-285 for any output size)."""
    lin = linear_image(accum)
    h, w = lin.shape[:2]
    x = np.arange(out_w, dtype=np.float64)
    y = np.arange(out_h, dtype=np.float64)
    qx = (x + 0.5) * (w / out_w) - 0.5
    qy = (y + 0.5) * (h / out_h) - 0.5
    ix, iy = np.floor(qx).astype(np.int64), np.floor(qy).astype(np.int64)
    fx, fy = qx - np.floor(qx), qy - np.floor(qy)
    out = np.zeros((out_h, out_w, 3), np.float64)
    for dx, wx in ((0, 1.0 - fx), (1, fx)):
        for dy, wy in ((0, 1.0 - fy), (1, fy)):
            xs = np.clip(ix + dx, 0, w - 1)
            ys = np.clip(iy + dy, 0, h - 1)
            out += (wy[:, None] * wx[None, :])[..., None] * lin[ys][:, xs]
    out = np.clip(out * exposure, 0.0, 1.0)
    if gamma:
        out = srgb_oetf(out)
    return out * 255.0


def worst_error(rgba8, accum, exposure=1.0, gamma=True):
    """max |8-bit value - expectation| over the RGB channels of a composite of any size."""
    oh, ow = rgba8.shape[:2]
    e = expected(accum, ow, oh, exposure, gamma)
    return float(np.abs(rgba8[..., :3].astype(np.float64) - e).max())


def probe_error(rgba8, accum):
    """verify.rs:239-249: same-size composite, five probe pixels of a 96 x 96 image."""
    h, w = accum.shape[:2]
    e = expected(accum, w, h).reshape(-1, 3)
    px = rgba8.reshape(-1, 4)[:, :3].astype(np.float64)
    idx = [0, 47 * w + 48, 48 * w + 48, 60 * w + 30, w * h - 1]
    return float(np.abs(px[idx] - e[idx]).max())


def gate_accum(w, h, seed):
    """An accumulation buffer that reaches every branch: counts 0 (max(count, 1)), 1 and many; linear values below
    the OETF's knee (0.0031308), inside [0, 1], above 1 (clamped) and exactly 0; smooth and sharp regions."""
    rng = np.random.default_rng(seed)
    n = rng.integers(0, 40, (h, w)).astype(np.float32)
    n[: h // 4, : w // 4] = 0.0
    n[h // 2:, : w // 8] = 1.0
    yy, xx = np.mgrid[0:h, 0:w]
    smooth = np.stack([0.5 + 0.5 * np.sin(xx / 7.0), 0.5 + 0.5 * np.cos(yy / 5.0), (xx + yy) / float(w + h)], -1)
    lin = smooth * rng.uniform(0.0, 1.3, (h, w, 1))
    lin[rng.random((h, w)) < 0.1] *= 0.002        # below the knee
    lin[rng.random((h, w)) < 0.05] = 0.0
    acc = np.zeros((h, w, 4), np.float32)
    acc[..., :3] = (lin * np.maximum(n, 1.0)[..., None]).astype(np.float32)
    acc[..., 3] = n
    return acc
