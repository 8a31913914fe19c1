"""Test scenes and parameter helpers shared by the tests (no GPU, no oracle).

Scenes restate crates/gps-render/src/verify.rs:15-39 (phase 2-4 scenes),
verify.rs:174-184 (70k splats) and verify.rs:774-789 (crowd of 3000).
"""
import math

import numpy as np

BG = (0.05, 0.05, 0.08)
GAUSS_DTYPE = np.dtype([("pos_opacity", "<f4", 4), ("log_scale", "<f4", 4), ("rot_wxyz", "<f4", 4), ("color", "<f4", 4)])


def pcg_py(v):
    m = 0xFFFFFFFF
    s = (v * 747796405 + 2891336453) & m
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & m
    return (w >> 22) ^ w


def quat_xyzw(x, y, z, w):
    n = math.sqrt(x * x + y * y + z * z + w * w)
    return (x / n, y / n, z / n, w / n)


def iso(z, o, col, x=0.0, s=0.15):
    return dict(pos=(x, 0.0, z), log_scale=(math.log(s),) * 3, opacity=o, color=col, rot_xyzw=(0.0, 0.0, 0.0, 1.0))


def phase_scenes():
    """verify.rs:21-39"""
    c = (0.9, 0.4, 0.2)
    return [
        ("single o=0.1", [iso(6.0, 0.1, c)]),
        ("single o=0.5", [iso(6.0, 0.5, c)]),
        ("single o=0.9", [iso(6.0, 0.9, c)]),
        ("front/back", [iso(9.0, 0.8, (0.15, 0.7, 0.25)), iso(5.0, 0.6, (0.9, 0.2, 0.2))]),
        ("anisotropic", [dict(pos=(0.0, 0.0, 6.0), log_scale=(math.log(0.28), math.log(0.10), math.log(0.06)),
                              rot_xyzw=quat_xyzw(0.20, -0.45, 0.30, 0.82), opacity=0.7, color=(0.3, 0.5, 0.95))]),
    ]


def many_splats(n=70_000):
    """verify.rs:174-184"""
    out = []
    for i in range(n):
        gx, gy = (i % 280) / 280.0 - 0.5, (i // 280) / 250.0 - 0.5
        out.append(dict(pos=(gx * 0.9, gy * 0.9, 6.0 + (i % 7) * 0.1), log_scale=(math.log(0.012),) * 3,
                        opacity=0.3 + 0.6 * ((i * 37 % 100) / 100.0),
                        color=((i % 5) / 5.0, (i % 3) / 3.0, 0.5), rot_xyzw=(0.0, 0.0, 0.0, 1.0)))
    return out


def crowd(n=3000, seed=4242):
    """verify.rs:774-789"""
    s = seed

    def rnd():
        nonlocal s
        s = pcg_py(s)
        return s / 4294967296.0

    out = []
    for _ in range(n):
        sc = 0.04 + 0.2 * rnd()
        pos = (-1.3 + 2.6 * rnd(), -1.3 + 2.6 * rnd(), 4.0 + 5.0 * rnd())
        ls = (math.log(sc), math.log(sc * (0.3 + rnd())), math.log(sc * 0.5))
        q = quat_xyzw(rnd() - 0.5, rnd() - 0.5, rnd() - 0.5, rnd() + 0.2)
        o = 0.2 + 0.7 * rnd()
        col = (rnd(), rnd(), rnd())
        out.append(dict(pos=pos, log_scale=ls, rot_xyzw=q, opacity=o, color=col))
    return out


def to_gpu(scene):
    """GpuGaussian::from_oracle (gps.rs:51-58): f32, rotation stored (w, x, y, z)."""
    a = np.zeros(len(scene), GAUSS_DTYPE)
    for i, g in enumerate(scene):
        x, y, z, w = g.get("rot_xyzw", (0.0, 0.0, 0.0, 1.0))
        a[i]["pos_opacity"] = (*g["pos"], g["opacity"])
        a[i]["log_scale"] = (*g["log_scale"], 0.0)
        a[i]["rot_wxyz"] = (w, x, y, z)
        a[i]["color"] = (*g["color"], 0.0)
    return a


def to_image(accum):
    """verify.rs:54-56: linear RGB = sum / max(count, 1), as f64"""
    a = accum.astype(np.float64)
    return a[..., :3] / np.maximum(a[..., 3:4], 1.0)
