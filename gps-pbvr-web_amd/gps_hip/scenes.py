"""Synthetic 3DGS scenes and cameras of the benchmark workloads.

The reference renders real PLY captures (crates/gps-verify/src/main.rs:106-148); the workloads here are
synthetic scenes of the same shape.  Every generator is seeded (numpy PCG64) and returns a GpuGaussian
array (gps.rs:26-38) plus, for SH degree > 0, the channel-major SH rest coefficients (gps.rs:687-691).

* ``small_uniform``: 4,096 Gaussians, positions U[-1,1]^3, log-scales N(-3.5, 0.5^2), uniform random
  unit quaternions, opacity sigmoid(N(0, 1.5^2)), SH0 colour U[0,1].  Camera at distance 3 looking at
  the origin, vertical FOV 60 deg.
* ``clustered``: N Gaussians, 80 % from 64 isotropic clusters (centres U[-5,5]^3, sigma U[0.2, 1.0]) and
  20 % on a noisy z = -2 plane (x, y U[-5,5], z noise N(0, 0.05^2)); log-scales N(-4, 0.7^2), opacity
  logit N(0, 2^2); SH degree 3 with DC colour U[0,1] and higher bands N(0, 0.05^2).  Orbit cameras
  around the origin with +Z up, yaw 0.6, pitch 0.25, vertical FOV 60 deg.
"""
import math

import numpy as np

from . import GAUSSIAN_DTYPE, SH_C0, look_at, orbit_camera

FOV_60 = math.radians(60.0)


def _random_quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q  # (w, x, y, z)


def _pack(pos, log_scale, quat_wxyz, opacity, color):
    n = pos.shape[0]
    g = np.zeros(n, GAUSSIAN_DTYPE)
    g["pos_opacity"][:, :3] = pos
    g["pos_opacity"][:, 3] = opacity
    g["log_scale"][:, :3] = log_scale
    g["rot_wxyz"] = quat_wxyz
    g["color"][:, :3] = color
    return g


def small_uniform(n=4096, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    ls = rng.normal(-3.5, 0.5, (n, 3))
    q = _random_quats(rng, n)
    op = 1.0 / (1.0 + np.exp(-rng.normal(0.0, 1.5, n)))
    col = rng.uniform(0.0, 1.0, (n, 3))
    return _pack(pos, ls, q, op, col)


def small_uniform_camera(w=256, h=256):
    return look_at((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, -1.0, 0.0), FOV_60, w, h)


def clustered(n=560_000, seed=1, sh_degree=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    n_plane = n // 5
    n_cl = n - n_plane
    centres = rng.uniform(-5.0, 5.0, (64, 3))
    sig = rng.uniform(0.2, 1.0, 64)
    which = rng.integers(0, 64, n_cl)
    p_cl = centres[which] + rng.standard_normal((n_cl, 3)) * sig[which, None]
    p_pl = np.stack([rng.uniform(-5.0, 5.0, n_plane), rng.uniform(-5.0, 5.0, n_plane),
                     -2.0 + rng.normal(0.0, 0.05, n_plane)], axis=1)
    pos = np.concatenate([p_cl, p_pl])
    perm = rng.permutation(n)
    pos = pos[perm]
    ls = rng.normal(-4.0, 0.7, (n, 3))
    q = _random_quats(rng, n)
    op = 1.0 / (1.0 + np.exp(-rng.normal(0.0, 2.0, n)))
    dc_col = rng.uniform(0.0, 1.0, (n, 3))
    g = _pack(pos, ls, q, op, dc_col)
    sh = None
    if sh_degree > 0:
        r = (sh_degree + 1) ** 2 - 1
        sh = rng.normal(0.0, 0.05, (n, 3, r)).astype(np.float32).reshape(-1)
    return g, sh


ORBIT = dict(near=2.0, mid=6.0, far=15.0)


def clustered_camera(view="near", w=1280, h=720):
    return orbit_camera((0.0, 0.0, 0.0), ORBIT[view], 0.6, 0.25, (0.0, 0.0, 1.0), FOV_60, w, h)


def dc_from_color(color):
    """Inverse of GpuGaussian::from_activated's colour (gps.rs:42): f_dc = (c - 0.5) / SH_C0."""
    return (np.asarray(color) - 0.5) / SH_C0
