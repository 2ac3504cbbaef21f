"""Sanitizer runs of the CPU-side code (the oracle and the library's host C++),
built here with gcc / g++ and -fsanitize=address,undefined: no device code is
involved.  Malformed image files must be rejected with an error code and no
memory error."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="gcc / g++ not available")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    assert r.returncode == 0, (cmd[0], r.stdout[-2000:], r.stderr[-4000:])
    return r


def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_san"
    _run(["gcc", *SAN, "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-ffp-contract=off",
          "-I", os.path.join(ROOT, "oracle"), "-o", str(exe),
          os.path.join(ROOT, "tests", "sanitize", "oracle_san_main.c"),
          os.path.join(ROOT, "oracle", "kamino_oracle.c"), "-lm"])
    r = _run([str(exe)], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert "clean" in r.stdout


def test_host_code_under_asan_ubsan(tmp_path):
    exe = tmp_path / "init_san"
    _run(["g++", *SAN, "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
          os.path.join(ROOT, "tests", "sanitize", "init_san_main.cpp"),
          os.path.join(ROOT, "kaminogpu_amd", "csrc", "kamino_init.cpp")])
    files = {
        "ok.ppm": b"P6\n3 2\n255\n" + bytes(range(18)),
        "ascii.pgm": b"P2\n# c\n2 2\n15\n0 5 10 15\n",
        "short.ppm": b"P6\n40 40\n255\n" + b"\x01" * 10,          # truncated pixels
        "huge.ppm": b"P6\n100000 100000\n255\n",                   # absurd size
        "neg.ppm": b"P6\n-3 2\n255\n",
        "maxval.ppm": b"P6\n1 1\n65535\n\x00\x00\x00\x00\x00\x00",  # 16-bit samples
        "magic.ppm": b"P9\n1 1\n255\n\x00",
        "empty.ppm": b"",
        "ascii_short.pgm": b"P2\n3 3\n255\n1 2\n",
    }
    for name, data in files.items():
        (tmp_path / name).write_bytes(data)
    args = [str(tmp_path / n) for n in files] + [str(tmp_path / "missing.ppm")]
    r = _run([str(exe), *args], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    out = r.stdout
    assert "clean" in out
    # the two valid files load (status 0), every malformed one is rejected
    for name in files:
        line = [l for l in out.splitlines() if l.startswith(str(tmp_path / name)) and "nTheta=4" in l][0]
        ok = name in ("ok.ppm", "ascii.pgm")
        assert (" density 0 " in line) == ok, line
