"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The PLY loader (gps-pbvr-web_amd/csrc/gps_ply.cpp, which parses untrusted files) and the oracle
(oracle/gps_core_f64.c, oracle/gps_mirror.c) are compiled with g++/gcc and -fsanitize=address,undefined into the small
drivers of tests/sanitize/ and run on the PLY fixture plus a few hundred corrupted files, and on small scenes through
every oracle renderer.  A sanitizer report fails the test, and the instrumented builds must compute exactly what the
uninstrumented libgps_hip.so and liboracle.so compute."""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from scenes import BG, phase_scenes, to_gpu

gh = pytest.importorskip("gps_hip")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "sanitize")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")


def _tool(name):
    if shutil.which(name) is None:
        pytest.skip(f"{name} is not available")
    return name


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=600, **kw)
    report = r.stderr
    assert "AddressSanitizer" not in report and "runtime error" not in report and "LeakSanitizer" not in report, \
        report[-4000:]
    assert r.returncode == 0, (r.returncode, report[-4000:])
    return r.stdout


@pytest.fixture(scope="module")
def ply_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("san") / "ply_check")
    _run([_tool("g++"), "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
          os.path.join(ROOT, "gps-pbvr-web_amd", "csrc", "gps_ply.cpp"), os.path.join(HERE, "ply_check.cpp"),
          "-o", out])
    return out


@pytest.fixture(scope="module")
def oracle_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("san") / "oracle_check")
    _run([_tool("gcc"), "-std=c11", *SAN, "-D_GNU_SOURCE", "-ffp-contract=off", "-fno-fast-math", "-mfma",
          "-fopenmp", os.path.join(ROOT, "oracle", "gps_core_f64.c"), os.path.join(ROOT, "oracle", "gps_mirror.c"),
          os.path.join(HERE, "oracle_check.c"), "-lm", "-o", out])
    return out


def _checksum(*arrays):
    """The drivers' checksum: sum of b[i] (i * 2654435761 + 1) over the arrays' bytes, modulo 2^64."""
    b = np.concatenate([np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8) for a in arrays]
                       + [np.zeros(0, np.uint8)]).astype(np.uint64)
    i = np.arange(b.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int((b * (i * np.uint64(2654435761) + np.uint64(1))).sum(dtype=np.uint64))


def _corrupted(raw, trials, seed):
    """Byte flips anywhere and in the header, truncations, deleted or duplicated header lines, random tails."""
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    lines = raw[:end].split(b"\n")
    rng = np.random.default_rng(seed)
    for t in range(trials):
        k = t % 5
        if k == 0:
            b = bytearray(raw)
            for _ in range(int(rng.integers(1, 8))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        elif k == 1:
            b = bytearray(raw)
            for _ in range(int(rng.integers(1, 4))):
                b[int(rng.integers(0, end))] = int(rng.integers(0, 256))
        elif k == 2:
            b = raw[:int(rng.integers(0, len(raw)))]
        elif k == 3:
            ls = list(lines)
            i = int(rng.integers(1, len(ls) - 2))
            ls = ls[:i] + ls[i + 1:] if rng.random() < 0.5 else ls[:i] + [ls[i]] + ls[i:]
            b = b"\n".join(ls) + raw[end:]
        else:
            b = raw[:int(rng.integers(0, end))] + rng.integers(0, 256, int(rng.integers(0, 200)), np.uint8).tobytes()
        yield bytes(b)


def test_ply_loader_under_sanitizers(ply_check, tmp_path):
    raw = open(os.path.join(ROOT, "tests", "golden", "synthetic_sh1.ply"), "rb").read()
    ascii_ = ("ply\nformat ascii 1.0\nelement vertex 2\n"
              + "".join(f"property float {p}\n" for p in ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                                          "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
                                                          "rot_3"])
              + "end_header\n" + "1 2 3 0.1 0.2 0.3 0.5 -1 -1 -1 1 0 0 0\n" * 2).encode()
    files = [raw, ascii_] + list(_corrupted(raw, 400, 7)) + list(_corrupted(ascii_, 100, 8))
    paths = []
    for i, b in enumerate(files):
        p = tmp_path / f"f{i:04d}.ply"
        p.write_bytes(b)
        paths.append(str(p))
    lines = _run([ply_check] + paths).strip().split("\n")
    assert len(lines) == len(files)
    outcomes = set()
    for b, line in zip(files, lines):
        st, count, deg, h = (int(x) for x in line.split())
        try:
            g = gh.load_ply(b)
        except gh.PlyError as e:
            assert st == e.code
            outcomes.add(e.kind)
            continue
        assert st == 0 and count == g.count and deg == g.sh_degree
        assert h == _checksum(g.position, g.sh_rest, g.to_gpu())
        outcomes.add("ok")
    assert "ok" in outcomes and len(outcomes) >= 4, outcomes


def _params(method, spp, path=1, **pb):
    p = O.MParams(spp_side=spp, density_scale=1.0, max_points_per_splat=4096, near=0.05, exposure=1.0, srgb_output=1,
                  sh_degree=2, path=path, method=method)
    p.background[:] = BG
    p.pbvr.calibration = pb.get("cal", 3)
    p.pbvr.radial_correction = pb.get("radial", 0)
    p.pbvr.low_pass_jitter = 1
    p.pbvr.base_k = 512.0
    p.pbvr.reference_pixel_length = 0.01
    return p


def test_oracle_under_sanitizers(oracle_check, tmp_path):
    scene = phase_scenes()[3][1] + phase_scenes()[4][1] + phase_scenes()[1][1]
    g = to_gpu(scene)
    n, deg, W, H, seed, count = len(scene), 2, 40, 32, 3, 2
    rng = np.random.default_rng(1)
    sh = rng.normal(0.0, 0.1, n * 3 * 8).astype(np.float32)
    cam = O.default_camera(120.0)
    sets = [_params(0, 2), _params(0, 1, path=0), _params(2, 1, radial=1), _params(4, 2), _params(1, 2, cal=1)]
    popt = O.default_particle_options(level=3, rule=1, radial_correction=1, per_particle_seed=1)
    mopt = O.default_mc_options(per_particle_seed=1)
    orc = O.gaussians_f64(scene)
    blob = struct.pack("<11I", n, deg, W, H, seed, count, len(sets), ctypes.sizeof(O.MParams),
                       ctypes.sizeof(O.OrcGaussian), ctypes.sizeof(popt), ctypes.sizeof(mopt))
    blob += g.tobytes() + sh.tobytes() + bytes(cam) + b"".join(bytes(p) for p in sets) + bytes(orc)
    blob += bytes(popt) + bytes(mopt)
    (tmp_path / "in.bin").write_bytes(blob)
    _run([oracle_check, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")])
    out = memoryview((tmp_path / "out.bin").read_bytes())
    at = 0

    def take(dtype, count_):
        nonlocal at
        a = np.frombuffer(out[at:at + np.dtype(dtype).itemsize * count_], dtype).copy()
        at += a.nbytes
        return a

    X = W * H
    for p in sets:
        S = X * p.spp_side ** 2
        rc = int(take(np.int32, 1)[0])
        got = dict(accum=take(np.float32, 4 * X), stats=take(np.uint64, 16), counts=take(np.uint32, n),
                   offsets=take(np.uint32, n), depth_id=take(np.uint64, S), total=int(take(np.uint32, 1)[0]),
                   skipped=int(take(np.uint32, 1)[0]))
        assert rc == 0
        ref = O.mirror_render(g, sh, deg, cam, p, W, H, seed, count, want_buffers=True)
        assert np.array_equal(got["accum"].view(np.uint32), ref["accum"].reshape(-1).view(np.uint32))
        assert np.array_equal(got["stats"], ref["stats"])
        for k in ("counts", "offsets", "depth_id"):
            assert np.array_equal(got[k], ref[k]), k
        assert got["total"] == ref["total"] and got["skipped"] == ref["skipped"]
    rc = int(take(np.int32, 1)[0])
    acc, stats = take(np.float32, 4 * X), take(np.uint64, 16)
    keys, vals, vis = take(np.uint32, n), take(np.uint32, n), int(take(np.uint32, 1)[0])
    rgba = take(np.uint8, 16 * X)
    ref = O.mirror_sorted(g, sh, deg, cam, sets[0], W, H)
    assert rc == 0 and vis == ref["visible"]
    assert np.array_equal(acc.view(np.uint32), ref["accum"].reshape(-1).view(np.uint32))
    assert np.array_equal(stats, ref["stats"]) and np.array_equal(keys, ref["keys"]) and np.array_equal(vals, ref["vals"])
    assert np.array_equal(rgba, O.mirror_composite(ref["accum"], 1.0, 1, 2 * W, 2 * H).reshape(-1))
    img, st = take(np.float64, 3 * X), take(np.float64, 3)
    ref, rst = O.render_particles_3d(scene, cam, W, H, 2, count, BG, seed, popt)
    assert np.array_equal(img, ref.reshape(-1)) and tuple(st) == (rst.points, rst.candidates, rst.subpixel_writes)
    img, st = take(np.float64, 3 * X), take(np.float64, 3)
    ref, rst = O.render_monte_carlo(scene, cam, W, H, 2, count, BG, seed, mopt)
    assert np.array_equal(img, ref.reshape(-1)) and tuple(st) == (rst.points, rst.candidates, rst.subpixel_writes)
    img = take(np.float64, 3 * X)
    assert np.array_equal(img, O.render_analytic(scene, cam, W, H, BG).reshape(-1))
    assert at == len(out)
