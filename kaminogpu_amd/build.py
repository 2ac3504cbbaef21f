"""Builds libkamino.so (gfx950) and the `kamino` driver in-tree with hipcc.

    python -m kaminogpu_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libkamino.so")
DRIVER = os.path.join(PKG, "bin", "kamino")

HIP_SOURCES = ["kamino_kernels.hip", "kamino_dft.hip", "kamino_runtime.hip", "kamino_engine.hip",
               "kamino_noalloc.hip", "kamino_init.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# gfx950 only; packed FP32 instructions are disabled for every HIP source.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return out


def build_library() -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str | None:
    src = os.path.join(PKG, "driver", "kamino_main.cpp")
    if not os.path.exists(src):
        return None
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", DRIVER, src,
          "-L", PKG, "-lkamino", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"])
    return DRIVER


def build() -> None:
    build_library()
    build_driver()


if __name__ == "__main__":
    build()
    print(LIB)
