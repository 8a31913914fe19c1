"""The sorted alpha path's per-quadrant culling (gps-pbvr-web_amd/csrc/gps_sorted.hip: box_touches, ellipse_touch)
against the blend's own per-pixel arithmetic, restated here in float32 (numpy rounds every operation correctly and
fuses nothing, as the kernel is compiled): a splat the tests reject for an 8x8 quadrant must have r2 > k^2 at all 64
pixel centres, so that its alpha is zero there and leaving it out changes no bit.  The hardware reciprocal and
reciprocal square root that ellipse_touch uses are modelled as the exact ones moved by up to one ulp either way.

Cases: random splats near and far from the quadrant, from sub-pixel to thousands of pixels, needles down to 1e-3 of
their major axis, every orientation; and adversarial ones whose k is set just above the smallest r2 over the
quadrant's pixels, so that exactly one pixel or a few keep the splat."""
import numpy as np

F = np.float32
QUAD = 8


def pixel_r2(a, b, qx0, qy0):
    """r2 at the quadrant's 64 pixel centres, as k_sp_blend computes it (splat_r2 with the staged denominators):
    arrays of cases x 64."""
    lx = np.arange(64) % QUAD
    ly = np.arange(64) // QUAD
    X = (qx0[:, None] + lx[None, :]).astype(F)
    Y = (qy0[:, None] + ly[None, :]).astype(F)
    qx, qy = X + F(0.5), Y + F(0.5)
    ax, ay = a[:, 0:1], a[:, 1:2]
    bx, by, bz, bw = b[:, 0:1], b[:, 1:2], b[:, 2:3], b[:, 3:4]
    den1 = bx * bx + by * by
    den2 = bz * bz + bw * bw
    dx, dy = qx - ax, qy - ay
    u = (dx * bx + dy * by) / den1
    v = (dx * bz + dy * bw) / den2
    return u * u + v * v


def one_ulp_off(x, rng):
    """x moved by -1, 0 or +1 ulp at random (a hardware approximation within one ulp)."""
    step = rng.integers(-1, 2, x.shape)
    up, dn = np.nextafter(x, F(np.inf)), np.nextafter(x, F(-np.inf))
    return np.where(step > 0, up, np.where(step < 0, dn, x)).astype(F)


def box_touches(a, b, qx0, qy0):
    ex = a[:, 3] * (np.abs(b[:, 0]) + np.abs(b[:, 2])) + F(1.0)
    ey = a[:, 3] * (np.abs(b[:, 1]) + np.abs(b[:, 3])) + F(1.0)
    return (~(np.floor(a[:, 0] - ex) > qx0 + F(QUAD - 1)) & ~(np.ceil(a[:, 0] + ex) < qx0) &
            ~(np.floor(a[:, 1] - ey) > qy0 + F(QUAD - 1)) & ~(np.ceil(a[:, 1] + ey) < qy0))


def edge_min(c, lo, hi, m_cc, m_tt, m_ct, r_tt, det):
    t = -m_ct * c * r_tt
    tol = F(2.0 ** -12) * (np.abs(lo) + np.abs(hi) + F(1.0))
    inner = ~(t < lo - tol) & ~(t > hi + tol)
    e = np.where(t < lo, lo, hi).astype(F)
    at_end = m_tt * e * e + F(2.0) * m_ct * e * c + m_cc * c * c
    return np.where(inner, (det * r_tt) * (c * c), at_end).astype(F)


def ellipse_touch(a, b, qx0, qy0, rng):
    with np.errstate(all="ignore"):
        bx, by, bz, bw = b[:, 0], b[:, 1], b[:, 2], b[:, 3]
        den1 = bx * bx + by * by
        den2 = bz * bz + bw * bw
        ok = ((den1 > F(2.0 ** -20)) & (den1 < F(2.0 ** 20)) & (den2 > F(2.0 ** -20)) & (den2 < F(2.0 ** 20)) &
              (np.abs(a[:, 0]) < F(2.0 ** 24)) & (np.abs(a[:, 1]) < F(2.0 ** 24)) & (a[:, 3] >= 0) &
              (a[:, 3] < F(2.0 ** 10)))
        x0 = qx0 + F(0.5) - a[:, 0]
        x1 = qx0 + F(QUAD - 0.5) - a[:, 0]
        y0 = qy0 + F(0.5) - a[:, 1]
        y1 = qy0 + F(QUAD - 0.5) - a[:, 1]
        centre_in = (x0 <= 0) & (x1 >= 0) & (y0 <= 0) & (y1 >= 0)
        i1 = one_ulp_off(F(1.0) / (den1 * den1), rng)
        i2 = one_ulp_off(F(1.0) / (den2 * den2), rng)
        mxx = bx * bx * i1 + bz * bz * i2
        mxy = bx * by * i1 + bz * bw * i2
        myy = by * by * i1 + bw * bw * i2
        pos = (mxx > 0) & (myy > 0)
        rxx = one_ulp_off(F(1.0) / mxx, rng)
        ryy = one_ulp_off(F(1.0) / myy, rng)
        det = np.maximum(mxx * myy - mxy * mxy, F(0.0))
        f = np.minimum(np.minimum(edge_min(y0, x0, x1, myy, mxx, mxy, rxx, det),
                                  edge_min(y1, x0, x1, myy, mxx, mxy, rxx, det)),
                       np.minimum(edge_min(x0, y0, y1, mxx, myy, mxy, ryy, det),
                                  edge_min(x1, y0, y1, mxx, myy, mxy, ryy, det)))
        xm = np.maximum(np.abs(x0), np.abs(x1))
        ym = np.maximum(np.abs(y0), np.abs(y1))
        f_lo = f - F(2.0 ** -15) * (mxx + myy + F(2.0) * np.abs(mxy)) * ((xm + ym) * (xm + ym))
        rs1 = one_ulp_off(F(1.0) / np.sqrt(den1), rng)
        rs2 = one_ulp_off(F(1.0) / np.sqrt(den2), rng)
        dist = F(2.0 ** -21) * F(1.01) * (xm + ym) * (rs1 + rs2)
        bound = a[:, 3] * F(1.0 + 2.0 ** -14) + dist
        reject = f_lo * F(1.0 - 2.0 ** -14) > bound * bound
        return ~ok | centre_in | ~pos | ~reject


def splats(rng, n, spread, smin, smax, thin, aligned=False):
    """n splats around their quadrants: records a = (x, y, o, k), b = (b1, b2) as k_sp_project writes them; aligned:
    axes along x and y, where the binning box is tight, half of them centred on a row of pixel centres."""
    qx0 = (rng.integers(0, 160, n) * QUAD).astype(F)
    qy0 = (rng.integers(0, 90, n) * QUAD).astype(F)
    cx = qx0 + F(4.0) + rng.normal(0.0, 1.0, n).astype(F) * spread
    cy = qy0 + F(4.0) + rng.normal(0.0, 1.0, n).astype(F) * spread
    if aligned:
        cy = np.where(rng.integers(0, 2, n) == 1, np.floor(cy) + F(0.5), cy).astype(F)
    s1 = np.exp(rng.uniform(np.log(smin), np.log(smax), n))
    s2 = s1 * np.exp(rng.uniform(np.log(thin), 0.0, n))
    th = rng.integers(0, 4, n) * (np.pi / 2.0) if aligned else rng.uniform(0.0, 2.0 * np.pi, n)
    c, s = np.cos(th), np.sin(th)
    b = np.stack([c * s1, s * s1, -s * s2, c * s2], 1).astype(F)
    k = rng.uniform(0.0, np.sqrt(2.0 * np.log(255.0)), n).astype(F)
    a = np.stack([cx, cy, np.full(n, 0.5, F), k], 1).astype(F)
    return a, b, qx0, qy0


def check(a, b, qx0, qy0, rng):
    r2 = pixel_r2(a, b, qx0, qy0)
    kk = (a[:, 3] * a[:, 3])[:, None]
    some_pixel = (~(r2 > kk)).any(axis=1)   # the kernel keeps alpha for !(r2 > k^2), NaN included
    ell = ellipse_touch(a, b, qx0, qy0, rng)
    box = box_touches(a, b, qx0, qy0)
    assert not (some_pixel & ~ell).any(), "the ellipse bound rejected a splat with a pixel inside"
    assert not (some_pixel & ~box).any(), "the binning box rejected a splat with a pixel inside"
    return ell, box, some_pixel


def test_cull_never_drops_a_contributing_splat():
    rng = np.random.default_rng(20261204)
    rejected = kept_inside = 0
    for spread, smin, smax, thin in ((6.0, 0.3, 8.0, 1e-3), (20.0, 0.3, 60.0, 1e-2), (300.0, 5.0, 3000.0, 1e-3),
                                     (3.0, 0.05, 2.0, 1.0)):
        a, b, qx0, qy0 = splats(rng, 200_000, spread, smin, smax, thin)
        ell, box, some = check(a, b, qx0, qy0, rng)
        rejected += int((~ell & box).sum())
        kept_inside += int(some.sum())
    # the bound does cull: a test that kept everything would pass the check above trivially
    assert rejected > 100_000 and kept_inside > 100_000


def test_cull_keeps_splats_reaching_one_pixel_by_an_ulp():
    """k just above the smallest r2 over the quadrant: one pixel (or a few) keeps a non-zero alpha, and both tests
    must keep the splat; with splats along the axes as well, where the binning box meets the ellipse."""
    rng = np.random.default_rng(7)
    for spread, smin, smax, thin, aligned in ((6.0, 0.3, 8.0, 1e-3, False), (30.0, 1.0, 40.0, 1e-2, False),
                                              (2.0, 0.3, 3.0, 1.0, False), (12.0, 0.3, 20.0, 1e-2, True)):
        a, b, qx0, qy0 = splats(rng, 200_000, spread, smin, smax, thin, aligned)
        m = pixel_r2(a, b, qx0, qy0).min(axis=1)
        keep = (m > 0) & (m < F(11.0))
        a, b, qx0, qy0, m = a[keep], b[keep], qx0[keep], qy0[keep], m[keep]
        k = np.sqrt(m).astype(F)
        for _ in range(3):   # the smallest float k with k * k >= m
            k = np.where(k * k < m, np.nextafter(k, F(np.inf)), k).astype(F)
        a[:, 3] = k
        ell, box, some = check(a, b, qx0, qy0, rng)
        assert some.all() and ell.all() and box.all()
