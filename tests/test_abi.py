"""CPU checks of the C ABI boundary: the library loads without a GPU, exports
every function include/kamino.h declares, and rejects bad arguments with
status codes (never by exiting)."""
import ctypes as C
import os
import re

import pytest

import kaminogpu_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "kamino.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kam_[a-z_]+)\s*\(", text)))


def test_header_declares_the_python_symbol_list():
    assert declared_functions() == sorted(K.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(K.LIB_PATH)
    missing = [s for s in declared_functions() if not hasattr(L, s)]
    assert not missing, missing


def test_bad_shapes_are_rejected_without_a_gpu():
    L = K.lib()
    h = C.c_void_p()
    for nt, nph, dt in [(3, 6, 0.1), (64, 100, 0.1), (64, 128, 0.0), (7 * 11, 154, 0.1), (4096, 8192, 0.1)]:
        d = K.SimDesc(nt, nph, 1.0, dt, 1, 0, 0)
        rc = L.kam_sim_create(C.byref(d), C.byref(h))
        assert rc != K.KAM_OK
        assert L.kam_last_error().decode()


def test_sim_bytes_is_28_per_cell_plus_watchdog_block():
    assert K.lib().kam_sim_bytes(256, 512) == 28 * 256 * 512 + 256


def test_no_device_in_this_container_is_reported_not_fatal():
    n = K.device_count()
    assert n >= 0
    if n == 0:
        with pytest.raises(K.KaminoError):
            K.Sim(64, 0.005)


def test_product_initial_state_matches_the_oracle_generator():
    """kam_init_fields (product) and ko_init_fields (oracle) are two independent
    implementations of the same seeded recipe: equal to FP32 rounding."""
    import numpy as np

    from oracle import oracle as O

    for nt, L, seed, dens in [(16, 4, 3, 0), (64, 8, 1, 1), (48, 12, 9, 0)]:
        got = K.init_fields(nt, L, seed, dens)
        ref = O.init_fields(nt, 2 * nt, L, seed, dens)
        for g, r in zip(got, ref):
            assert g.dtype == np.float32
            np.testing.assert_allclose(g, r, rtol=2e-6, atol=1e-6 * max(1.0, np.abs(r).max()))


def test_tuning_table_round_trip_without_a_gpu():
    default = K.tune_get(192)
    K.tune_set(192, 12, 64, 4)
    assert K.tune_get(192) == (12, 64, 4)
    K.tune_set(192, 0, 0, 0)
    assert K.tune_get(192) == default
    with pytest.raises(K.KaminoError):
        K.tune_set(2, 1, 1, 1)


def test_noalloc_step_rejects_bad_arguments_without_a_gpu():
    L = K.lib()
    d = K.SimDesc(64, 128, 1.0, 0.01, 1, 0, 0)
    assert L.kam_sim_step_noalloc(None, 1, 1, None, 1, None) == -1  # KAM_EINVAL
    assert L.kam_sim_step_noalloc(C.byref(d), None, 1, None, 1, None) == -1
    assert L.kam_sim_step_noalloc(C.byref(d), 1, 1, None, -1, None) == -1
    bad = K.SimDesc(64, 100, 1.0, 0.01, 1, 0, 0)
    assert L.kam_sim_step_noalloc(C.byref(bad), 1, 1, None, 1, None) == -1
    assert L.kam_last_error().decode()


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of the header's structs have the C compiler's size
    and field offsets (a field added to a struct on one side only would shift
    every later field)."""
    import subprocess

    from kaminogpu_amd import engine as E

    mirrors = {"kam_sim_desc": K.SimDesc, "kam_engine_desc": E.EngineDesc, "kam_req_status": E.ReqStatus,
               "kam_stats": E.EngineStats}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "kamino.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in filter(None, out):
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
