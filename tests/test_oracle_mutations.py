"""The oracle's closed-form pins catch plausible mistakes.

Each case builds a copy of oracle/kamino_oracle.c with one deliberate mistake
(a flipped sign, a dropped average, a missing 1/R, a wrong metric update)
into a temporary directory, and runs the pin tests that cover that line
against it in a child process (KAMINO_ORACLE_LIB points the oracle wrapper at
the broken build).  The pins must fail on every broken copy, and the intact
sources must pass them (checked first, with the same build recipe).  This runs
on the CPU only; nothing here touches the product library.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "kamino_oracle.c")
PINS = "tests/test_oracle.py"

# (name, exact source text, replacement, pin tests that must fail)
MUTATIONS = [
    ("theta departure sign", "const double pTheta = gTheta - deltaTheta;",
     "const double pTheta = gTheta + deltaTheta;",
     ["test_uniform_flow_departure_point", "test_linear_meridional_flow_rk2_rule"]),
    ("phi departure sign", "const double pPhi = gPhi - deltaPhi;",
     "const double pPhi = gPhi + deltaPhi;", ["test_uniform_flow_departure_point"]),
    ("plain midpoint instead of the RK2 average", "const double averuTheta = 0.5 * (muTheta + guTheta);",
     "const double averuTheta = muTheta;", ["test_linear_meridional_flow_rk2_rule"]),
    ("theta metric without 1/R", "const double cofTheta = dt / radius;", "const double cofTheta = dt;",
     ["test_uniform_flow_departure_point", "test_linear_meridional_flow_rk2_rule"]),
    ("phi metric without sin(theta)", "const double cofPhi = dt / latRadius;",
     "const double cofPhi = dt / radius;", ["test_uniform_flow_departure_point"]),
    ("geometric G without 1/R", "const double G = dt * cos(gTheta) / (radius * sin(gTheta));",
     "const double G = dt * cos(gTheta) / sin(gTheta);", ["test_geometric_zonal_flow_at_dt"]),
    ("geometric v' sign", "const double vNext = vPrev + G * uNext * uNext;",
     "const double vNext = vPrev - G * uNext * uNext;", ["test_geometric_zonal_flow_at_dt"]),
]


def _build(tmp, text):
    src = os.path.join(tmp, "kamino_oracle.c")
    with open(src, "w") as f:
        f.write(text)
    lib = os.path.join(tmp, "libkamino_oracle.so")
    subprocess.run(["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c11",
                    "-D_POSIX_C_SOURCE=200809L", "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", lib,
                    src, "-lm"], check=True)
    return lib


def _run_pins(lib, tests):
    env = dict(os.environ, KAMINO_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", PINS,
                        "-k", " or ".join(tests)], cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode, r.stdout


@pytest.fixture(scope="module")
def source():
    with open(SRC) as f:
        return f.read()


def test_intact_sources_pass_the_pins(source, tmp_path):
    lib = _build(str(tmp_path), source)
    tests = sorted({t for m in MUTATIONS for t in m[3]})
    rc, out = _run_pins(lib, tests)
    assert rc == 0, out


@pytest.mark.parametrize("name,old,new,tests", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_mutation_is_caught(source, tmp_path, name, old, new, tests):
    assert source.count(old) == 1, f"mutation site for '{name}' not found exactly once"
    lib = _build(str(tmp_path), source.replace(old, new))
    for t in tests:
        rc, out = _run_pins(lib, [t])
        assert rc != 0, f"'{name}' passed {t}:\n{out}"
