"""Trajectory export host logic (no GPU): key interpolation and orbit keys as web/index.html:448-457 define them,
the PNG encoding, and the key -> camera mapping."""
import math
import os
import struct
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gps-pbvr-web_amd"))
from gps_hip import trajectory as T  # noqa: E402

K0 = {"target": [0.0, 0.0, 0.0], "distance": 4.0, "yaw": 0.0, "pitch": 0.1, "fov_y": 0.9, "up_y": -1}
K1 = {"target": [1.0, 2.0, 3.0], "distance": 8.0, "yaw": 1.0, "pitch": 0.3, "fov_y": 0.5, "up_y": -1}
K2 = {"target": [1.0, 0.0, 1.0], "distance": 6.0, "yaw": 3.0, "pitch": -0.2, "fov_y": 0.7, "up_y": -1}


def test_cam_at_is_piecewise_linear():
    keys = [K0, K1, K2]
    assert T.cam_at(keys, 0.0) == K0 and T.cam_at(keys, 1.0)["target"] == K2["target"]
    m = T.cam_at(keys, 0.25)   # halfway between K0 and K1
    assert m["target"] == [0.5, 1.0, 1.5] and m["distance"] == 6.0 and m["yaw"] == 0.5
    assert abs(m["pitch"] - 0.2) < 1e-15 and abs(m["fov_y"] - 0.7) < 1e-15 and m["up_y"] == -1
    m = T.cam_at(keys, 0.75)   # halfway between K1 and K2
    assert m["target"] == [1.0, 1.0, 2.0] and m["distance"] == 7.0 and m["yaw"] == 2.0
    assert T.cam_at([K1], 0.6) == K1


def test_orbit_keys_turn_once():
    ks = T.orbit_keys(K1)
    assert [k["yaw"] - K1["yaw"] for k in ks] == [0.0, 0.5 * math.pi, math.pi, 1.5 * math.pi, 2.0 * math.pi]
    assert all(k["distance"] == K1["distance"] and k["target"] == K1["target"] for k in ks)


def test_png_encoding_round_trips():
    img = (np.arange(7 * 5 * 4) % 251).astype(np.uint8).reshape(5, 7, 4)
    png = T.png_rgba8(img)
    assert png[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(png):
        n = struct.unpack(">I", png[pos:pos + 4])[0]
        tag, data = png[pos + 4:pos + 8], png[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", png[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(tag + data) & 0xFFFFFFFF
        chunks[tag] = data
        pos += 12 + n
    assert struct.unpack(">IIBBBBB", chunks[b"IHDR"]) == (7, 5, 8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(5, 1 + 7 * 4)
    assert (rows[:, 0] == 0).all() and np.array_equal(rows[:, 1:].reshape(5, 7, 4), img)


def test_key_camera_sits_at_the_key_distance():
    cam = T.key_camera(K1, 64, 48)
    eye = np.array(cam.view_pos)
    assert abs(np.linalg.norm(eye - np.array(K1["target"])) - K1["distance"]) < 1e-9
    steep = T.key_camera(dict(K1, pitch=5.0), 64, 48)   # pitch is clamped away from the pole
    assert np.all(np.isfinite(np.array(steep.view_rot)))
