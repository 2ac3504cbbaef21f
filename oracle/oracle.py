"""ctypes view of the FP64 oracle (oracle/kamino_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, the smoke entry point and the
CPU-baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.  The product package (kaminogpu_amd) does not import it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KAMINO_ORACLE_LIB names another build of the same sources (the mutation tests
# load deliberately broken copies this way to show that the pins catch them).
LIB_PATH = os.environ.get("KAMINO_ORACLE_LIB") or os.path.join(HERE, "_build", "libkamino_oracle.so")

VPHI, VTHETA, CENTERED = 0, 1, 2
PHASE_ADVECT, PHASE_GEOMETRIC, PHASE_PROJECT, PHASE_ALL = 1, 2, 4, 7

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class Stats(C.Structure):
    _fields_ = [
        ("pole_escapes", C.c_int64),
        ("ambiguous_uphi", C.c_int64),
        ("ambiguous_utheta", C.c_int64),
        ("ambiguous_rho", C.c_int64),
    ]


def build() -> str:
    """Compile the oracle with its own Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ko_validate_coord.restype = C.c_double
        L.ko_validate_coord.argtypes = [_dp, _dp]
        L.ko_sample.restype = C.c_double
        L.ko_sample.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.POINTER(C.c_int64)]
        L.ko_solve_cubic.restype = C.c_double
        L.ko_solve_cubic.argtypes = [C.c_double] * 3
        L.ko_poisson_rows.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ko_thomas.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ko_cyclic_reduction.restype = C.c_int
        L.ko_cyclic_reduction.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.ko_fft.restype = C.c_int
        L.ko_fft.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.ko_splitmix64.restype = C.c_uint64
        L.ko_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.ko_uniform.restype = C.c_double
        L.ko_uniform.argtypes = [C.POINTER(C.c_uint64)]
        L.ko_normal.restype = C.c_double
        L.ko_normal.argtypes = [C.POINTER(C.c_uint64)]
        L.ko_real_sh.restype = C.c_double
        L.ko_real_sh.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.ko_init_fields.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64, C.c_int,
                                     _dp, _dp, _dp]
        L.ko_advect.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp,
                                _dp, _dp, _dp, _u8p, _u8p, _u8p, C.c_double, C.POINTER(Stats)]
        L.ko_advect_particles.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp,
                                          C.c_int64, _dp, _dp, _u8p, C.c_double]
        L.ko_geometric.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.ko_geometric_mask.argtypes = [C.c_int, C.c_int, _u8p, _u8p, _u8p, _u8p]
        L.ko_divergence.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.ko_project.restype = C.c_int
        L.ko_project.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.ko_step.restype = C.c_int
        L.ko_step.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp,
                              C.POINTER(Stats)]
        L.ko_run.restype = C.c_double
        L.ko_run.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                             _dp, _dp, _dp]
        L.ko_max_threads.restype = C.c_int
        L.ko_fbm.restype = C.c_float
        L.ko_fbm.argtypes = [C.c_float, C.c_float]
        L.ko_fbm_hash.restype = C.c_float
        L.ko_fbm_hash.argtypes = [C.c_float, C.c_float]
        _fp32 = C.POINTER(C.c_float)
        L.ko_init_fbm.argtypes = [C.c_int, C.c_int, C.c_double, _fp32, _fp32]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _pu8(a):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)


def empty_fields(n_theta: int, n_phi: int):
    return (np.zeros((n_theta, n_phi)), np.zeros((n_theta - 1, n_phi)), np.zeros((n_theta, n_phi)))


def init_fields(n_theta, n_phi, L, seed, density_kind, radius=1.0):
    """Seeded curl-of-streamfunction velocity and density (see kamino_oracle.h)."""
    up, ut, rho = empty_fields(n_theta, n_phi)
    lib().ko_init_fields(n_theta, n_phi, radius, L, seed, density_kind, _p(up), _p(ut), _p(rho))
    return up, ut, rho


def sample(kind, field, n_theta, n_phi, phi, theta):
    esc = C.c_int64(0)
    f = np.ascontiguousarray(field, dtype=np.float64)
    v = lib().ko_sample(kind, _p(f), n_theta, n_phi, np.pi / n_theta, phi, theta, C.byref(esc))
    return v, esc.value


def validate_coord(phi, theta):
    a, b = C.c_double(phi), C.c_double(theta)
    r = lib().ko_validate_coord(C.byref(a), C.byref(b))
    return a.value, b.value, r


def solve_cubic(a, b, c):
    return lib().ko_solve_cubic(a, b, c)


def poisson_rows(n_theta, n_phi, n):
    a, b, c = (np.zeros(n_theta) for _ in range(3))
    lib().ko_poisson_rows(n_theta, n_phi, n, _p(a), _p(b), _p(c))
    return a, b, c


def thomas(a, b, c, d):
    x = np.zeros_like(d)
    lib().ko_thomas(len(d), _p(a), _p(b), _p(c), _p(np.ascontiguousarray(d)), _p(x))
    return x


def cyclic_reduction(a, b, c, d):
    x = np.zeros_like(d)
    rc = lib().ko_cyclic_reduction(len(d), _p(a), _p(b), _p(c), _p(np.ascontiguousarray(d)), _p(x))
    if rc != 0:
        raise ValueError("cyclic reduction needs a power-of-two size")
    return x


def fft(x, sign):
    re = np.ascontiguousarray(np.real(x), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64).copy()
    if lib().ko_fft(len(re), _p(re), _p(im), sign) != 0:
        raise ValueError("bad fft arguments")
    return re + 1j * im


def advect(n_theta, n_phi, dt, u_phi, u_theta, rho, radius=1.0, tol=1e-5, stats=None):
    """Returns (u_phi*, u_theta*, rho*, amb_phi, amb_theta, amb_rho, stats)."""
    u_phi = np.ascontiguousarray(u_phi, dtype=np.float64)
    u_theta = np.ascontiguousarray(u_theta, dtype=np.float64)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    op, ot, orr = empty_fields(n_theta, n_phi)
    ap = np.zeros(op.shape, np.uint8)
    at = np.zeros(ot.shape, np.uint8)
    ar = np.zeros(orr.shape, np.uint8)
    st = stats if stats is not None else Stats()
    lib().ko_advect(n_theta, n_phi, radius, dt, _p(u_phi), _p(u_theta), _p(rho), _p(op), _p(ot),
                    _p(orr), _pu8(ap), _pu8(at), _pu8(ar), tol, C.byref(st))
    return op, ot, orr, ap, at, ar, st


def advect_particles(n_theta, n_phi, dt, u_phi, u_theta, pos, radius=1.0, tol=1e-5):
    """One particle step; pos is (n, 2) (phi, theta).  Returns (pos*, amb)."""
    u_phi = np.ascontiguousarray(u_phi, dtype=np.float64)
    u_theta = np.ascontiguousarray(u_theta, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
    out = np.empty_like(pos)
    amb = np.zeros(len(pos), np.uint8)
    lib().ko_advect_particles(n_theta, n_phi, radius, dt, _p(u_phi), _p(u_theta), len(pos),
                              _p(pos), _p(out), _pu8(amb), tol)
    return out, amb


def geometric(n_theta, n_phi, dt, u_phi, u_theta, radius=1.0):
    up = np.array(u_phi, dtype=np.float64, order="C")
    ut = np.array(u_theta, dtype=np.float64, order="C")
    lib().ko_geometric(n_theta, n_phi, radius, dt, _p(up), _p(ut))
    return up, ut


def geometric_mask(n_theta, n_phi, amb_phi, amb_theta):
    op = np.zeros((n_theta, n_phi), np.uint8)
    ot = np.zeros((n_theta - 1, n_phi), np.uint8)
    lib().ko_geometric_mask(n_theta, n_phi, _pu8(np.ascontiguousarray(amb_phi, np.uint8)),
                            _pu8(np.ascontiguousarray(amb_theta, np.uint8)), _pu8(op), _pu8(ot))
    return op, ot


def divergence(n_theta, n_phi, u_phi, u_theta):
    d = np.zeros((n_theta, n_phi))
    lib().ko_divergence(n_theta, n_phi, _p(np.ascontiguousarray(u_phi, np.float64)),
                        _p(np.ascontiguousarray(u_theta, np.float64)), _p(d))
    return d


def project(n_theta, n_phi, u_phi, u_theta, solver=0):
    """Returns (u_phi, u_theta, spec[nPhi][nTheta] complex in the reference's
    shifted mode order, p[nTheta][nPhi])."""
    up = np.array(u_phi, dtype=np.float64, order="C")
    ut = np.array(u_theta, dtype=np.float64, order="C")
    sr = np.zeros((n_phi, n_theta))
    si = np.zeros((n_phi, n_theta))
    p = np.zeros((n_theta, n_phi))
    rc = lib().ko_project(n_theta, n_phi, _p(up), _p(ut), _p(sr), _p(si), _p(p), solver)
    if rc != 0:
        raise ValueError("projection failed (solver/size)")
    return up, ut, sr + 1j * si, p


def step(n_theta, n_phi, dt, u_phi, u_theta, rho, phases=PHASE_ALL, radius=1.0):
    up = np.array(u_phi, dtype=np.float64, order="C")
    ut = np.array(u_theta, dtype=np.float64, order="C")
    rr = np.array(rho, dtype=np.float64, order="C")
    st = Stats()
    rc = lib().ko_step(n_theta, n_phi, radius, dt, phases, _p(up), _p(ut), _p(rr), C.byref(st))
    if rc != 0:
        raise ValueError("step failed")
    return up, ut, rr, st


def run(n_theta, n_phi, dt, u_phi, u_theta, rho, steps, threads=0, phases=PHASE_ALL, radius=1.0):
    """Advances in place; returns wall seconds of the stepping loop."""
    return lib().ko_run(n_theta, n_phi, radius, dt, phases, steps, threads, _p(u_phi), _p(u_theta),
                        _p(rho))


def max_threads() -> int:
    return lib().ko_max_threads()


def fbm(x, y):
    return float(lib().ko_fbm(float(x), float(y)))


def fbm_hash(x, y):
    return float(lib().ko_fbm_hash(float(x), float(y)))


def init_fbm(n_theta, n_phi, radius=1.0):
    """FP32 (u_phi, u_theta) of the reference's FBM initializer."""
    up = np.empty((n_theta, n_phi), np.float32)
    ut = np.empty((n_theta - 1, n_phi), np.float32)
    fp = C.POINTER(C.c_float)
    lib().ko_init_fbm(n_theta, n_phi, radius, up.ctypes.data_as(fp), ut.ctypes.data_as(fp))
    return up, ut


def rel_l2(a, b, mask=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        keep = ~np.asarray(mask, bool)
        a, b = a[keep], b[keep]
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
