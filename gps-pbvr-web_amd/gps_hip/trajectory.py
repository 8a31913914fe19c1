"""Trajectory export: the viewer's orbit / keyframe export (web/index.html:389-391, 433-469) on the batch API.

Camera keys have the viewer's `camera_json` form (crates/gps-web/src/lib.rs:386-392): target, distance, yaw,
pitch, fov_y and up_y (+1 or -1).  A trajectory of `count` frames interpolates the keys piecewise linearly
(`cam_at`, index.html:448-453); `orbit_keys` makes one full turn around a key (index.html:454-457).  Every frame
accumulates the same number of ensembles; the frames are recorded as one batch and launched as one hipGraph
(the viewer renders them one frame loop at a time).  The export is a ZIP with frame_NNNN.png, stats.csv (one row
of counters per frame, CSV_COLS) and session.json (the session with the trajectory's keys).

    python -m gps_hip.trajectory --ply scene.ply --frames 24 --ensembles 32 --out trajectory.zip
"""
import argparse
import io
import json
import math
import struct
import sys
import zipfile
import zlib

import numpy as np

from . import Batch, GpsRenderer, Method, RenderParams, load_ply, orbit_camera, robust_sphere, frame_sphere_distance

# index.html:389
CSV_COLS = ["method", "seed", "accumulated", "points", "candidates", "visible", "truncated_splats", "dropped_points",
            "orphan", "undercovered", "width", "height", "spp", "yaw", "pitch", "distance"]
PITCH_LIMIT = 1.553   # crates/gps-core/src/camera.rs:28


def cam_at(keys, u):
    """Key at u in [0, 1] along the piecewise-linear path through `keys` (index.html:448-453)."""
    if len(keys) == 1:
        return dict(keys[0])
    x = u * (len(keys) - 1)
    i = min(int(math.floor(x)), len(keys) - 2)
    t = x - i
    a, b = keys[i], keys[i + 1]
    lerp = lambda p, q: p + (q - p) * t   # noqa: E731
    out = dict(a)
    out.update(target=[lerp(p, q) for p, q in zip(a["target"], b["target"])], distance=lerp(a["distance"], b["distance"]),
               yaw=lerp(a["yaw"], b["yaw"]), pitch=lerp(a["pitch"], b["pitch"]), fov_y=lerp(a["fov_y"], b["fov_y"]))
    return out


def orbit_keys(key):
    """One full turn around `key` (index.html:454-457)."""
    return [dict(key, yaw=key["yaw"] + f * 2.0 * math.pi) for f in (0.0, 0.25, 0.5, 0.75, 1.0)]


def key_camera(key, w, h):
    """The OracleCamera of a key (OrbitCamera::to_camera, camera.rs:63-65; pitch clamped as set_camera does)."""
    up = (0.0, -1.0, 0.0) if key.get("up_y", 1) < 0 else (0.0, 1.0, 0.0)
    pitch = max(-PITCH_LIMIT, min(PITCH_LIMIT, key["pitch"]))
    return orbit_camera(tuple(key["target"]), key["distance"], key["yaw"], pitch, up, key["fov_y"], w, h)


def png_rgba8(img):
    """PNG bytes of an (h, w, 4) uint8 image (8-bit RGBA, no filtering)."""
    h, w = img.shape[:2]
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def render_trajectory(renderer, params, keys, count, ensembles, method_label, first_seed=1, session=None,
                      use_graph=True):
    """Renders `count` frames along `keys` with `ensembles` ensembles each; returns (zip bytes, rows, timing)."""
    w, h = renderer.size()
    frames = [cam_at(keys, 0.0 if count == 1 else i / (count - 1)) for i in range(count)]
    cams = [key_camera(k, w, h) for k in frames]
    b = Batch()
    try:
        b.render([(renderer, cam, params, first_seed, ensembles) for cam in cams], use_graph=use_graph)
        rows = [",".join(["frame"] + CSV_COLS)]
        buf = io.BytesIO()
        with zipfile.ZipFile(buf, "w", zipfile.ZIP_STORED) as z:
            for i, k in enumerate(frames):
                st = b.read_stats(i)
                row = dict(method=method_label, seed=first_seed, accumulated=st.ensembles - st.skipped_ensembles,
                           points=st.points, candidates=st.candidates, visible=st.visible_splats,
                           truncated_splats=st.truncated_splats, dropped_points=st.dropped_points,
                           orphan=st.orphan_subpixels, undercovered=st.undercovered_splats, width=w, height=h,
                           spp=params.spp_side * params.spp_side, yaw=k["yaw"], pitch=k["pitch"],
                           distance=k["distance"])
                rows.append(",".join([str(i)] + [str(row[c]).replace(",", " ").replace('"', " ") for c in CSV_COLS]))
                z.writestr(f"frame_{i:04d}.png", png_rgba8(b.composite(i, params)))
            z.writestr("stats.csv", "\n".join(rows) + "\n")
            sess = dict(session or {"format": "gps-pbvr-session", "version": 1})
            sess["trajectory"] = keys
            z.writestr("session.json", json.dumps(sess, indent=2))
        return buf.getvalue(), rows, b.timing()
    finally:
        b.close()


METHODS = {"gps": (Method.Gps, 3, 0), "pbvr-ext-c3r": (Method.Extinction, 3, 1),
           "occupancy": (Method.ScreenOccupancy, 0, 0)}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--ply", required=True)
    ap.add_argument("--out", default="trajectory.zip")
    ap.add_argument("--frames", type=int, default=24)
    ap.add_argument("--ensembles", type=int, default=32)
    ap.add_argument("--size", default="640x360")
    ap.add_argument("--spp", type=int, default=1)
    ap.add_argument("--method", default="pbvr-ext-c3r", choices=sorted(METHODS))
    a = ap.parse_args(argv)
    w, h = (int(x) for x in a.size.split("x"))
    g = load_ply(a.ply)
    r = GpsRenderer(w, h, a.spp)
    r.set_gaussians_sh(g.to_gpu(), g.sh_rest, g.sh_degree) if g.sh_degree else r.set_gaussians(g.to_gpu())
    c, rad = robust_sphere(g.position)
    key = {"target": [float(x) for x in c], "distance": frame_sphere_distance(rad), "yaw": 0.6, "pitch": 0.25,
           "fov_y": 0.9, "up_y": -1}
    m, cal, radial = METHODS[a.method]
    p = RenderParams(spp_side=a.spp, method=m, max_points_per_splat=16384,
                     pbvr=dict(calibration=cal, radial_correction=radial))
    data, rows, t = render_trajectory(r, p, orbit_keys(key), max(1, min(720, a.frames)), a.ensembles, a.method,
                                      session={"format": "gps-pbvr-session", "version": 1,
                                               "scene": {"name": a.ply, "gaussians": g.count}, "camera": key})
    with open(a.out, "wb") as fh:
        fh.write(data)
    print(f"{a.out}: {a.frames} frames x {a.ensembles} ensembles in {t['gpu_ms']:.1f} GPU ms "
          f"({t['graph_nodes']} graph nodes)", file=sys.stderr)
    r.close()


if __name__ == "__main__":
    main()
