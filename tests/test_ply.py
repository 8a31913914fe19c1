"""3DGS PLY ingest (host code; no GPU): the gps-io loader's own tests (crates/gps-io/src/lib.rs:375-458) restated,
the reference's fixture tests/fixtures/synthetic_sh1.ply (copied to tests/golden/ as data) parsed independently
with numpy, the activation of crates/gps-render/src/gps.rs:41-49, and the framing helpers of
crates/gps-core/src/camera.rs:70-90."""
import math
import os
import struct
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gps-pbvr-web_amd"))
import gps_hip as gh  # noqa: E402

FIXTURE = os.path.join(ROOT, "tests", "golden", "synthetic_sh1.ply")
C0 = 0.28209479177387814


def header(fmt, n, with_sh1):
    """lib.rs:381-394"""
    h = f"ply\nformat {fmt} 1.0\nelement vertex {n}\n"
    for p in ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"]:
        h += f"property float {p}\n"
    if with_sh1:
        for i in range(9):
            h += f"property float f_rest_{i}\n"
    for p in ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]:
        h += f"property float {p}\n"
    return h + "end_header\n"


def err(data):
    with pytest.raises(gh.PlyError) as e:
        gh.load_ply(data)
    return e.value


def test_binary_roundtrip_sh1():
    """lib.rs:397-411"""
    b = bytearray(header("binary_little_endian", 2, True).encode())
    for v in range(2):
        for k in range(6 + 9 + 8):
            b += struct.pack("<f", (v * 100 + k) * 0.01)
    g = gh.load_ply(bytes(b))
    assert (g.count, g.sh_degree, g.binary) == (2, 1, True)
    assert g.position[1].tolist() == [np.float32(1.0), np.float32(1.01), np.float32(1.02)]
    assert g.sh_rest.size == 2 * 9
    assert abs(g.opacity_logit[0] - 0.15) < 1e-6   # opacity follows f_rest_8: index 6 + 9
    assert abs(g.rot_wxyz[0][0] - 0.19) < 1e-6


def test_ascii_and_activation():
    """lib.rs:413-421: sigmoid(0) = 0.5, identity rotation, colour = f_dc * C0 + 0.5"""
    s = header("ascii", 1, False) + "1 2 3  0.5 0 0  0  0 0 0  1 0 0 0\n"
    g = gh.load_ply(s.encode())
    assert not g.binary
    gg = g.to_gpu()[0]
    assert gg["pos_opacity"].tolist() == [1.0, 2.0, 3.0, 0.5]
    assert gg["rot_wxyz"].tolist() == [1.0, 0.0, 0.0, 0.0]
    assert gg["color"][0] == np.float32(0.5 * C0 + 0.5) and gg["color"][1] == np.float32(0.5)


def test_rejects_bad_input_without_panicking():
    """lib.rs:423-438"""
    e = err(b"hello")
    assert e.kind == "BadHeader" and str(e) == "bad PLY header: no end_header"
    assert err(b"nope\nend_header\n").kind == "NotPly"
    b = header("binary_little_endian", 5, False).encode() + bytes(10)
    e = err(b)
    assert e.kind == "Truncated" and str(e) == "file truncated: need 280 bytes of vertex data, have 10"
    e = err(header("binary_little_endian", gh.PLY_MAX_VERTICES + 1, False).encode())
    assert e.kind == "TooManyVertices"
    assert err((header("ascii", 1, False) + "nan 0 0 0 0 0 0 0 0 0 0 0 0 0\n").encode()).kind == "NonFinite"
    e = err(b"ply\nformat ascii 1.0\nelement vertex 0\nproperty float x\nend_header\n")
    assert e.kind == "MissingProperty" and str(e) == "missing required property 'y'"


def test_huge_declared_count_rejected_before_allocation():
    """lib.rs:440-446"""
    for fmt in ("binary_little_endian", "ascii"):
        assert err(header(fmt, gh.PLY_MAX_VERTICES, False).encode() + b"1 2 3\n").kind == "Truncated"


def test_rejects_bad_sh_count():
    """lib.rs:448-457"""
    h = "ply\nformat ascii 1.0\nelement vertex 0\n"
    for p in ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2", "rot_0",
              "rot_1", "rot_2", "rot_3", "f_rest_0"]:
        h += f"property float {p}\n"
    e = err((h + "end_header\n").encode())
    assert e.kind == "BadShCount" and str(e) == "1 f_rest_* properties is not 9, 24 or 45"


def test_header_errors_and_formats():
    """parse_header (lib.rs:148-227): format, element order, list properties, types, CRLF."""
    assert err(b"ply\nformat binary_big_endian 1.0\nelement vertex 0\nend_header\n").kind == "UnsupportedFormat"
    assert err(b"ply\nformat ascii 1.0\nelement face 1\nelement vertex 0\nend_header\n").kind == "UnsupportedFormat"
    assert err(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty list uchar int i\nend_header\n").kind == \
        "UnsupportedFormat"
    assert err(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty half x\nend_header\n").kind == "BadHeader"
    assert err(b"ply\nelement vertex 0\nend_header\n").kind == "BadHeader"
    assert err(b"ply\nformat ascii 1.0\nend_header\n").kind == "BadHeader"
    assert err(b"ply\nformat ascii 1.0\nelement vertex 0\nend_header").kind == "BadHeader"
    crlf = header("ascii", 1, False).replace("\n", "\r\n") + "1 2 3 0 0 0 0 0 0 0 1 0 0 0\r\n"
    assert gh.load_ply(crlf.encode()).position[0].tolist() == [1.0, 2.0, 3.0]
    # 13 values: long enough for the body-length pre-check (2 bytes per property), short one value
    bad = header("ascii", 1, False) + "1.0 2.0 3.0 0 0 0 0 0 0 0 1 0 0\n"
    e = err(bad.encode())
    assert e.kind == "BadAscii" and str(e) == "ASCII parse error at vertex 0: expected 14 values, got 13"
    assert err((header("ascii", 1, False) + "1 2 0x3 0 0 0 0 0 0 0 1 0 0 0\n").encode()).kind == "BadAscii"
    # a body shorter than 2 bytes per property is truncated before any token is read (lib.rs:285-293)
    assert err((header("ascii", 1, False) + "1 2 3 0 0 0 0 0 0 0 1 0 0\n").encode()).kind == "Truncated"


def test_scalar_types_widen_through_f64():
    """read_le (lib.rs:127-139): every PLY scalar type, little endian, converted to f32 through f64."""
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2", "rot_0",
             "rot_1", "rot_2", "rot_3"]
    types = [("char", "<b", -5), ("uchar", "<B", 200), ("short", "<h", -30000), ("ushort", "<H", 60000),
             ("int", "<i", -2_000_000_000), ("uint", "<I", 4_000_000_000), ("double", "<d", 0.1),
             ("float64", "<d", 1e-300), ("int8", "<b", 7), ("uint8", "<B", 1), ("int16", "<h", 2), ("uint16", "<H", 3),
             ("int32", "<i", 4), ("uint32", "<I", 5)]
    h = "ply\nformat binary_little_endian 1.0\nelement vertex 1\n"
    body = b""
    for name, (ty, fmt, v) in zip(names, types):
        h += f"property {ty} {name}\n"
        body += struct.pack(fmt, v)
    g = gh.load_ply((h + "end_header\n").encode() + body)
    got = np.concatenate([g.position[0], g.f_dc[0], [g.opacity_logit[0]], g.log_scale[0], g.rot_wxyz[0]])
    want = np.array([float(t[2]) for t in types], np.float64).astype(np.float32)
    assert np.array_equal(got, want)


def fixture_records():
    """The fixture's vertex records through numpy alone: the header's property list as a record dtype."""
    raw = open(FIXTURE, "rb").read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    names = [ln.split()[2].decode() for ln in raw[:end].split(b"\n") if ln.startswith(b"property float")]
    n = int([ln for ln in raw[:end].split(b"\n") if ln.startswith(b"element vertex")][0].split()[2])
    rec = np.frombuffer(raw[end:], np.dtype([(nm, "<f4") for nm in names]), count=n)
    return rec, names


def test_fixture_parses_like_numpy():
    rec, names = fixture_records()
    g = gh.load_ply(FIXTURE)
    assert (g.count, g.sh_degree) == (400, 1) and "nx" in names
    assert np.array_equal(g.position, np.stack([rec["x"], rec["y"], rec["z"]], 1))
    assert np.array_equal(g.f_dc, np.stack([rec[f"f_dc_{k}"] for k in range(3)], 1))
    assert np.array_equal(g.opacity_logit, rec["opacity"])
    assert np.array_equal(g.log_scale, np.stack([rec[f"scale_{k}"] for k in range(3)], 1))
    assert np.array_equal(g.rot_wxyz, np.stack([rec[f"rot_{k}"] for k in range(4)], 1))
    assert np.array_equal(g.sh_rest.reshape(400, 9), np.stack([rec[f"f_rest_{k}"] for k in range(9)], 1))


def test_fixture_activation_matches_the_generator():
    """tests/gen_synthetic_ply.py: 220 ground splats (opacity 0.9, scales 0.22 / 0.02 / 0.22, y = 0.5, identity
    rotation) and 3 x 60 blob splats (opacity 0.8, unit rotations, one colour per blob)."""
    g = gh.load_ply(FIXTURE)
    gg = g.to_gpu()
    op = gg["pos_opacity"][:, 3].astype(np.float64)
    assert np.abs(op[:220] - 0.9).max() < 1e-6 and np.abs(op[220:] - 0.8).max() < 1e-6
    assert np.allclose(np.exp(gg["log_scale"][:220, :3].astype(np.float64)), [0.22, 0.02, 0.22], rtol=1e-6)
    assert np.all(gg["pos_opacity"][:220, 1] == np.float32(0.5))
    assert np.all(gg["rot_wxyz"][:220] == np.array([1, 0, 0, 0], np.float32))
    q = gg["rot_wxyz"][220:].astype(np.float64)
    assert np.abs(np.linalg.norm(q, axis=1) - 1.0).max() < 1e-6
    cols = gg["color"][:, :3].astype(np.float64)
    ground = {(0.75, 0.75, 0.7), (0.35, 0.4, 0.5)}
    for c in cols[:220]:
        assert min(np.abs(c - np.array(k)).max() for k in ground) < 1e-6
    for b, col in enumerate([(0.9, 0.25, 0.2), (0.2, 0.7, 0.3), (0.25, 0.4, 0.95)]):
        assert np.abs(cols[220 + 60 * b:280 + 60 * b] - np.array(col)).max() < 1e-6
    # the exact f64 formulas, rounded once to f32
    want = (1.0 / (1.0 + np.exp(-g.opacity_logit.astype(np.float64)))).astype(np.float32)
    assert np.array_equal(gg["pos_opacity"][:, 3], want)
    assert np.array_equal(gg["color"][:, :3], (g.f_dc.astype(np.float64) * C0 + 0.5).astype(np.float32))


def test_robust_sphere_and_framing():
    """camera.rs:70-90"""
    assert gh.robust_sphere(np.zeros((0, 3))) is None
    pts = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [100, 0, 0]], np.float64)
    c, r = gh.robust_sphere(pts)
    assert c.tolist() == [2.0, 0.0, 0.0] and r == 98.0   # medians at index n/2; sorted distances [0,1,2,98], index 3
    g = gh.load_ply(FIXTURE)
    c, r = gh.robust_sphere(g.position)
    p = g.position.astype(np.float64)
    assert all(c[a] == np.sort(p[:, a])[200] for a in range(3))
    d = np.sort(np.linalg.norm(p - c, axis=1))
    assert r == d[300]
    assert abs(gh.frame_sphere_distance(1.0, 0.9) - 1.25 / math.sin(0.45)) < 1e-12


def test_fuzzed_inputs_return_errors_not_crashes():
    """lib.rs:8 (no panics on bad input): thousands of corrupted fixtures (byte flips in header and body,
    truncations, duplicated or deleted header lines, random bytes) either load or raise PlyError."""
    raw = open(FIXTURE, "rb").read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    rng = np.random.default_rng(2024)
    lines = raw[:end].split(b"\n")
    kinds = {}
    for trial in range(3000):
        k = trial % 6
        if k == 0:      # byte flips anywhere
            b = bytearray(raw)
            for _ in range(int(rng.integers(1, 8))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        elif k == 1:    # flips inside the header
            b = bytearray(raw)
            for _ in range(int(rng.integers(1, 4))):
                b[int(rng.integers(0, end))] = int(rng.integers(0, 256))
        elif k == 2:    # truncation
            b = raw[:int(rng.integers(0, len(raw)))]
        elif k == 3:    # delete or duplicate header lines
            ls = list(lines)
            i = int(rng.integers(1, len(ls) - 2))
            ls = ls[:i] + ls[i + 1:] if rng.random() < 0.5 else ls[:i] + [ls[i]] + ls[i:]
            b = b"\n".join(ls) + raw[end:]
        elif k == 4:    # random bytes after a valid prefix
            b = raw[:int(rng.integers(0, end))] + rng.integers(0, 256, int(rng.integers(0, 200)), np.uint8).tobytes()
        else:           # ASCII bodies with junk tokens
            n = int(rng.integers(0, 4))
            body = "".join(" ".join(rng.choice(["1", "-0.5", "nan", "inf", "1e40", "x", "", "1e", ".", "+2"],
                                               size=int(rng.integers(0, 30)))) + "\n" for _ in range(n))
            b = (f"ply\nformat ascii 1.0\nelement vertex {int(rng.integers(0, 4))}\n" +
                 "".join(f"property float {p}\n" for p in ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                                             "scale_0", "scale_1", "scale_2", "rot_0", "rot_1",
                                                             "rot_2", "rot_3"]) + "end_header\n" + body).encode()
        try:
            g = gh.load_ply(bytes(b))
            assert g.count <= gh.PLY_MAX_VERTICES and np.isfinite(g.position).all()
            kinds["ok"] = kinds.get("ok", 0) + 1
        except gh.PlyError as e:
            kinds[e.kind] = kinds.get(e.kind, 0) + 1
    assert len(kinds) >= 6, kinds   # many distinct outcomes were exercised
