"""The configKamino.txt driver (kaminogpu_amd/bin/kamino): CPU checks of its
config handling, and GPU checks that its frame schedule and outputs match the
library run directly."""
import os
import struct
import subprocess

import numpy as np
import pytest

import kaminogpu_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "kaminogpu_amd", "bin", "kamino")


def run(args, cwd):
    return subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=120)


def write_cfg(path, tokens):
    with open(path, "w") as f:
        f.write(" ".join(str(t) for t in tokens) + "\n")


def frame_steps(dt, DT, frames):
    """Steps per frame of Kamino::run (KaminoCore.cu:886-904), float32 clock."""
    T = np.float32(0.0)
    out = []
    for i in range(1, frames + 1):
        n = 0
        while T < np.float32(i) * np.float32(DT):
            n += 1
            T = np.float32(T + np.float32(dt))
        out.append(n + 1)
        T = np.float32(np.float32(i) * np.float32(DT))
    return out


def test_config_errors_are_reported(tmp_path):
    write_cfg(tmp_path / "short.txt", [1.0, 32, 200])
    r = run([str(tmp_path / "short.txt")], tmp_path)
    assert r.returncode == 1 and "16 tokens" in r.stderr
    write_cfg(tmp_path / "bad.txt", [1.0, 2, 200, 0.005, 0.02, 1, 0, 1, 1, 1, 1, "a", "b", "null", "null", "null"])
    r = run([str(tmp_path / "bad.txt")], tmp_path)
    assert r.returncode == 1 and "nTheta" in r.stderr
    r = run([], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stdout


def test_no_device_is_an_error_not_a_crash(tmp_path):
    if K.device_count() > 0:
        pytest.skip("a device is present")
    write_cfg(tmp_path / "c.txt", [1.0, 32, 200, 0.005, 0.02, 1, 0, 1, 1, 1, 1,
                                   str(tmp_path / "f"), str(tmp_path / "p"), "null", "null", "null"])
    r = run([str(tmp_path / "c.txt")], tmp_path)
    assert r.returncode == 2 and "device" in r.stderr


@pytest.mark.gpu
def test_driver_frames_match_library_run(tmp_path):
    nt, dt, DT, frames = 32, 0.005, 0.0123, 3
    write_cfg(tmp_path / "c.txt", [1.0, nt, 200, dt, DT, frames, 0, 1, 1, 1, 1,
                                   str(tmp_path / "frame"), str(tmp_path / "part"), "null", "null",
                                   "null", "seed=5", "L=6", "density=noise", "format=raw", "init=sh"])
    r = run([str(tmp_path / "c.txt")], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Frame 3 is ready" in r.stdout
    steps = frame_steps(dt, DT, frames)
    st = K.init_fields(nt, 6, 5, K.DENSITY_NOISE)
    # the config's dt is read as a C float, as the reference's main.cu:12 does
    # particleDensity 200 reaches initParticlesfromPic as an integer (KaminoSolver.cu:279)
    parts = K.particles_seed(nt, 200, 5)
    with K.Sim(nt, float(np.float32(dt))) as s:
        s.set_fields(*st)
        s.set_particles(parts)
        for i, n in enumerate(steps, start=1):
            s.step(n)
            ref = s.get_fields()
            rawp = open(tmp_path / f"part{i}.kamp", "rb").read()
            assert rawp[:4] == b"KAMP"
            (cnt,) = struct.unpack("<q", rawp[4:12])
            assert cnt == len(parts)
            np.testing.assert_array_equal(np.frombuffer(rawp[12:], np.float32).reshape(-1, 2),
                                          s.get_particles())
            raw = open(tmp_path / f"frame{i}.kamr", "rb").read()
            assert raw[:4] == b"KAMR"
            a, b, k = struct.unpack("<iiq", raw[4:20])
            assert (a, b, k) == (nt, 2 * nt, sum(steps[:i]))
            data = np.frombuffer(raw[28:], np.float32)
            N = nt * 2 * nt
            np.testing.assert_array_equal(data[:N].reshape(nt, 2 * nt), ref[0])
            np.testing.assert_array_equal(data[N:2 * N - 2 * nt].reshape(nt - 1, 2 * nt), ref[1])
            np.testing.assert_array_equal(data[2 * N - 2 * nt:].reshape(nt, 2 * nt), ref[2])


@pytest.mark.gpu
def test_driver_default_init_is_the_reference_fbm(tmp_path):
    nt, dt, DT = 32, 0.0005, 0.001
    write_cfg(tmp_path / "c.txt", [1.0, nt, 4, dt, DT, 1, 0, 1, 1, 1, 1,
                                   str(tmp_path / "frame"), str(tmp_path / "part"), "null", "null",
                                   "null", "format=raw", "particles=0"])
    r = run([str(tmp_path / "c.txt")], tmp_path)
    assert r.returncode == 0, r.stderr
    up, ut = K.init_fbm(nt)
    _, _, rho = K.init_fields(nt, 8, 1, K.DENSITY_BAND)
    raw0 = np.frombuffer(open(tmp_path / "frame0.kamr", "rb").read()[28:], np.float32)
    N = nt * 2 * nt
    np.testing.assert_array_equal(raw0[:N].reshape(nt, 2 * nt), up)
    np.testing.assert_array_equal(raw0[N:2 * N - 2 * nt].reshape(nt - 1, 2 * nt), ut)
    with K.Sim(nt, float(np.float32(dt))) as s:
        s.set_fields(up, ut, rho)
        s.step(sum(frame_steps(dt, DT, 1)))
        ref = s.get_fields()
    raw1 = np.frombuffer(open(tmp_path / "frame1.kamr", "rb").read()[28:], np.float32)
    np.testing.assert_array_equal(raw1[:N].reshape(nt, 2 * nt), ref[0])


@pytest.mark.gpu
def test_driver_image_inputs(tmp_path):
    """densityImage sets the initial density (initDensityfromPic); colorImage
    colours the particles, written as colorBGR[3i+1..3i+3] like
    write_particles_bgeo (KaminoSolver.cu:382-384)."""
    nt = 16
    rng = np.random.default_rng(4)
    rgb = rng.integers(0, 256, (10, 20, 3), dtype=np.uint8)
    (tmp_path / "d.ppm").write_bytes(b"P6\n20 10\n255\n" + rgb.tobytes())
    write_cfg(tmp_path / "c.txt", [1.0, nt, 1, 0.001, 0.001, 1, 0, 1, 1, 1, 1,
                                   str(tmp_path / "frame"), str(tmp_path / "part"),
                                   str(tmp_path / "d.ppm"), "null", str(tmp_path / "d.ppm"),
                                   "format=raw"])
    r = run([str(tmp_path / "c.txt")], tmp_path)
    assert r.returncode == 0, r.stderr
    raw0 = np.frombuffer(open(tmp_path / "frame0.kamr", "rb").read()[28:], np.float32)
    N = nt * 2 * nt
    np.testing.assert_array_equal(raw0[2 * N - 2 * nt:].reshape(nt, 2 * nt), K.image_density(tmp_path / "d.ppm", nt))
    write_cfg(tmp_path / "c2.txt", [1.0, nt, 1, 0.001, 0.001, 1, 0, 1, 1, 1, 1,
                                    str(tmp_path / "frame"), str(tmp_path / "part"),
                                    "null", "null", str(tmp_path / "d.ppm")])
    r = run([str(tmp_path / "c2.txt")], tmp_path)
    assert r.returncode == 0, r.stderr
    parts = K.particles_seed(nt, 1, 1)
    cols = K.image_colors(tmp_path / "d.ppm", nt, parts).reshape(-1)
    want = np.concatenate([cols, [0.0]])[1:].reshape(-1, 3)
    b = open(tmp_path / "part0.bgeo", "rb").read()
    pts = _check_bgeo(b, len(parts), [("v", 3), ("color", 3)], radius=1.0)
    np.testing.assert_array_equal(pts[:, 7:10], want)


@pytest.mark.gpu
def test_driver_bgeo_layout(tmp_path):
    nt = 16
    write_cfg(tmp_path / "c.txt", [2.0, nt, 200, 0.01, 0.02, 1, 0, 1, 1, 1, 1,
                                   str(tmp_path / "frame"), str(tmp_path / "part"), "null", "null",
                                   "null"])
    r = run([str(tmp_path / "c.txt")], tmp_path)
    assert r.returncode == 0, r.stderr
    _check_bgeo(open(tmp_path / "frame1.bgeo", "rb").read(), nt * 2 * nt, [("v", 3), ("density", 1)])
    npart = len(K.particles_seed(nt, 200, 1))
    _check_bgeo(open(tmp_path / "part1.bgeo", "rb").read(), npart, [("v", 3), ("color", 3)])


def _check_bgeo(b, n, attrs, radius=2.0):
    assert b[:5] == b"BgeoV"
    ver, npts, nprim, _, _, npattr = struct.unpack(">6i", b[5:29])
    assert (ver, npts, nprim, npattr) == (5, n, 1, len(attrs))
    off = 29 + 12  # vertex / primitive / detail attribute counts
    names = []
    for _ in range(npattr):
        (ln,) = struct.unpack(">H", b[off:off + 2])
        name = b[off + 2:off + 2 + ln].decode()
        size, _typ = struct.unpack(">Hi", b[off + 2 + ln:off + 8 + ln])
        names.append((name, size))
        off += 8 + ln + 4 * size
    assert names == attrs
    width = 4 + sum(k for _, k in attrs)
    pts = np.frombuffer(b[off:off + npts * width * 4], ">f4").reshape(npts, width)
    np.testing.assert_allclose(np.linalg.norm(pts[:, :3], axis=1), radius, rtol=1e-6)  # on the sphere
    np.testing.assert_allclose(pts[:, 3], 1.0)
    assert b[-2:] == b"\x00\xff"
    return pts


@pytest.mark.gpu
def test_driver_bgeo_velocity_is_map_v_to_sphere(tmp_path):
    """The grid .bgeo's v attribute of frame 0 against the closed form of
    write_data_bgeo (KaminoSolver.cu:305-353) and mapVToSphere (:411-423),
    evaluated here in numpy from the same seeded state: the cell-centred
    (u_theta, u_phi) = ((v_N + v_S) / 2, (u_W + u_E) / 2), with
    jNorth = jSouth = 0 in row 0 and nTheta - 2 in the last row, mapped to
    (cos t cos p u_t - sin p u_p, -sin t u_t, cos t sin p u_t + cos p u_p) at
    the cell centre (p, t) = (i h, (j + 1/2) h)."""
    nt, R = 16, 2.0
    write_cfg(tmp_path / "c.txt", [R, nt, 4, 0.01, 0.02, 0, 0, 1, 1, 1, 1,
                                   str(tmp_path / "frame"), str(tmp_path / "part"), "null", "null",
                                   "null", "init=sh", "seed=3", "L=4", "density=noise"])
    r = run([str(tmp_path / "c.txt")], tmp_path)
    assert r.returncode == 0, r.stderr
    up, ut, rho = K.init_fields(nt, 4, 3, K.DENSITY_NOISE, radius=R)
    up, ut = up.astype(np.float64), ut.astype(np.float64)
    h = np.pi / nt
    uph = 0.5 * (up + np.roll(up, -1, 1))
    jn = np.r_[0, np.arange(nt - 2), nt - 2]
    js = np.r_[0, np.arange(1, nt - 1), nt - 2]
    uth = 0.5 * (ut[jn] + ut[js])
    t = ((np.arange(nt) + 0.5) * h)[:, None]
    p = (np.arange(2 * nt) * h)[None, :]
    v = np.stack([np.cos(t) * np.cos(p) * uth - np.sin(p) * uph, -np.sin(t) * uth + 0 * p,
                  np.cos(t) * np.sin(p) * uth + np.cos(p) * uph], -1).reshape(-1, 3)
    pts = _check_bgeo(open(tmp_path / "frame0.bgeo", "rb").read(), nt * 2 * nt, [("v", 3), ("density", 1)],
                      radius=R)
    scale = np.abs(v).max()
    assert scale > 1e-3
    np.testing.assert_allclose(pts[:, 4:7], v, rtol=0, atol=2e-6 * scale)
    np.testing.assert_array_equal(pts[:, 7], rho.reshape(-1))
    # the pole rows use one u_theta row (jNorth = jSouth), not an average with a neighbour
    row0 = pts[:2 * nt, 4:7]
    np.testing.assert_allclose(row0[:, 1], -np.sin(h / 2) * ut[0], rtol=1e-5)


@pytest.mark.gpu
def test_driver_resume_continues_the_run(tmp_path):
    """N frames straight equal k frames, then a run resumed from frame k's raw
    snapshot (fields and particles) for the remaining N - k frames, bitwise;
    the resumed run writes frames k+1 .. N and leaves frame k alone."""
    nt, dt, DT, N, k = 32, 0.005, 0.0123, 4, 2
    common = ["seed=5", "L=6", "density=noise", "format=raw", "init=sh"]

    def cfg(name, frames, prefix):
        write_cfg(tmp_path / name, [1.0, nt, 50, dt, DT, frames, 0, 1, 1, 1, 1,
                                    str(tmp_path / (prefix + "f")), str(tmp_path / (prefix + "p")),
                                    "null", "null", "null", *common])
        return str(tmp_path / name)

    r = run([cfg("straight.txt", N, "s")], tmp_path)
    assert r.returncode == 0, r.stderr
    r = run([cfg("first.txt", k, "a")], tmp_path)
    assert r.returncode == 0, r.stderr
    snap = (tmp_path / f"af{k}.kamr").read_bytes()
    r = run([cfg("rest.txt", N, "a"), f"resume={tmp_path / f'af{k}.kamr'}",
             f"resume_particles={tmp_path / f'ap{k}.kamp'}"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert f"Frame {k} is ready" not in r.stdout and f"Frame {N} is ready" in r.stdout
    assert (tmp_path / f"af{k}.kamr").read_bytes() == snap
    for i in range(k + 1, N + 1):
        assert (tmp_path / f"af{i}.kamr").read_bytes() == (tmp_path / f"sf{i}.kamr").read_bytes()
        assert (tmp_path / f"ap{i}.kamp").read_bytes() == (tmp_path / f"sp{i}.kamp").read_bytes()
    # a snapshot that is not at a frame time of this config is refused
    write_cfg(tmp_path / "bad2.txt", [1.0, nt, 50, dt, 0.0071, N, 0, 1, 1, 1, 1, str(tmp_path / "cf"),
                                      str(tmp_path / "cp"), "null", "null", "null", *common])
    r = run([str(tmp_path / "bad2.txt"), f"resume={tmp_path / f'af{k}.kamr'}"], tmp_path)
    assert r.returncode == 2 and "frame time" in r.stderr
