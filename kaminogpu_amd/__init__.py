"""kaminogpu_amd -- MI355X-native KaminoGPU time step.

Python view of libkamino.so (include/kamino.h) for tests, benchmarks and
scripting.  The compute path is the HIP library; if it is not built, importing
`lib()` raises: there is no CPU fallback.

The class `KaminoSolver` mirrors the reference's solver API
(KaminoSolver.cuh:99-111): construct with the grid, set the fields,
`stepForward`, read the fields back.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkamino.so")

KAM_OK = 0
PHASE_ADVECT, PHASE_DIVFFT, PHASE_TRI, PHASE_IFFT = 1, 8, 16, 32
PHASE_PROJECT = 8 | 16 | 32
PHASE_ALL = 1 | 8 | 16 | 32
PREC_FP32 = 0
PREC_FP64_PRESSURE = 1  # FP64 pressure spectrum, inverse FFT and gradient

# every symbol include/kamino.h declares
ABI_SYMBOLS = [
    "kam_last_error", "kam_version", "kam_device_count",
    "kam_sim_create", "kam_sim_destroy", "kam_sim_set_fields", "kam_sim_get_fields",
    "kam_sim_step", "kam_sim_step_phases", "kam_sim_get_spectrum", "kam_sim_sync", "kam_sim_status",
    "kam_sim_bytes", "kam_sim_benchmark", "kam_init_fields", "kam_engine_completed",
    "kam_sim_set_particles", "kam_sim_get_particles", "kam_sim_particle_count", "kam_particles_seed",
    "kam_init_fbm", "kam_sim_poisson_residual", "kam_image_density", "kam_image_colors",
    "kam_tune_set", "kam_tune_get", "kam_autotune", "kam_tune_load",
    "kam_engine_create", "kam_engine_destroy", "kam_engine_admit", "kam_engine_step",
    "kam_engine_poll", "kam_engine_fetch", "kam_engine_cancel", "kam_engine_sync",
    "kam_engine_stats", "kam_engine_reset_stats", "kam_engine_profile", "kam_engine_now",
    "kam_sim_step_noalloc", "kam_sim_get_pressure", "kam_tune_set_transform", "kam_tune_get_transform",
]
TRANSFORM_FFT, TRANSFORM_DFT, TRANSFORM_FFT8 = 0, 1, 2
NOALLOC_SCRATCH_BYTES = 16 << 20  # KAM_NOALLOC_SCRATCH_BYTES


class KaminoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kamino error {code}: {msg}")
        self.code = code


class SimDesc(C.Structure):
    _fields_ = [
        ("n_theta", C.c_int), ("n_phi", C.c_int), ("radius", C.c_double), ("dt", C.c_double),
        ("geometric", C.c_int), ("precision", C.c_int), ("device", C.c_int),
    ]


_fp = C.POINTER(C.c_float)
_lib = None


def lib():
    """Loads libkamino.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python -m kaminogpu_amd.build")
    L = C.CDLL(LIB_PATH)
    L.kam_last_error.restype = C.c_char_p
    L.kam_version.restype = C.c_char_p
    L.kam_device_count.restype = C.c_int
    L.kam_sim_create.argtypes = [C.POINTER(SimDesc), C.POINTER(C.c_void_p)]
    L.kam_sim_destroy.argtypes = [C.c_void_p]
    L.kam_sim_set_fields.argtypes = [C.c_void_p, _fp, _fp, _fp]
    L.kam_sim_get_fields.argtypes = [C.c_void_p, _fp, _fp, _fp]
    L.kam_sim_step.argtypes = [C.c_void_p, C.c_int]
    L.kam_sim_step_phases.argtypes = [C.c_void_p, C.c_int]
    L.kam_sim_get_spectrum.argtypes = [C.c_void_p, _fp]
    L.kam_sim_get_pressure.argtypes = [C.c_void_p, _fp]
    L.kam_sim_sync.argtypes = [C.c_void_p]
    L.kam_sim_status.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    _dpp = C.POINTER(C.c_double)
    L.kam_sim_benchmark.argtypes = [C.c_void_p, C.c_int, C.c_int, _dpp, _dpp, _dpp]
    L.kam_sim_bytes.restype = C.c_size_t
    L.kam_sim_bytes.argtypes = [C.c_int, C.c_int]
    L.kam_tune_set.argtypes = [C.c_int] * 4
    _ip = C.POINTER(C.c_int)
    L.kam_tune_get.argtypes = [C.c_int, _ip, _ip, _ip]
    L.kam_autotune.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, _dpp]
    L.kam_tune_load.argtypes = [C.c_char_p]
    L.kam_tune_set_transform.argtypes = [C.c_int, C.c_int]
    L.kam_tune_get_transform.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.kam_init_fields.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64, C.c_int, _fp, _fp, _fp]
    L.kam_sim_set_particles.argtypes = [C.c_void_p, _fp, C.c_int64]
    L.kam_sim_get_particles.argtypes = [C.c_void_p, _fp, C.c_int64]
    L.kam_sim_particle_count.restype = C.c_int64
    L.kam_sim_particle_count.argtypes = [C.c_void_p]
    L.kam_particles_seed.argtypes = [C.c_int, C.c_double, C.c_uint64, _fp, C.POINTER(C.c_int64)]
    L.kam_init_fbm.argtypes = [C.c_int, C.c_int, C.c_double, _fp, _fp]
    L.kam_sim_poisson_residual.argtypes = [C.c_void_p, _dpp, _dpp, _dpp]
    L.kam_image_density.argtypes = [C.c_char_p, C.c_int, C.c_int, _fp]
    L.kam_image_colors.argtypes = [C.c_char_p, C.c_int, C.c_int, _fp, C.c_int64, _fp]
    L.kam_sim_step_noalloc.argtypes = [C.POINTER(SimDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_void_p]
    _bind_engine(L)
    _lib = L
    return L


def _bind_engine(L):
    from . import engine as _engine

    _engine.bind(L)


def check(rc):
    if rc != KAM_OK:
        raise KaminoError(rc, lib().kam_last_error().decode())
    return rc


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def device_count() -> int:
    return lib().kam_device_count()


def tune_get(n_theta):
    r, w, f = C.c_int(), C.c_int(), C.c_int()
    check(lib().kam_tune_get(int(n_theta), C.byref(r), C.byref(w), C.byref(f)))
    return r.value, w.value, f.value


def tune_set(n_theta, adv_rows=0, adv_cols=0, fft_rows=0):
    check(lib().kam_tune_set(int(n_theta), int(adv_rows), int(adv_cols), int(fft_rows)))


def tune_set_transform(n_theta, transform):
    """Phi transform for simulations of n_theta created from now on (TRANSFORM_FFT / _FFT8 / _DFT)."""
    check(lib().kam_tune_set_transform(int(n_theta), int(transform)))


def tune_get_transform(n_theta):
    t = C.c_int()
    check(lib().kam_tune_get_transform(int(n_theta), C.byref(t)))
    return t.value


def autotune(n_theta, batch, steps=10, cache_path=None, device=0):
    """Times candidate tile shapes on the device and keeps the fastest;
    returns the best per-kernel times (ms)."""
    k = (C.c_double * 4)()
    check(lib().kam_autotune(int(device), int(n_theta), int(batch), int(steps),
                             cache_path.encode() if cache_path else None, k))
    return np.array(list(k))


def tune_load(cache_path):
    """Loads a tuning cache written by autotune; returns the n_theta values it holds."""
    check(lib().kam_tune_load(cache_path.encode()))
    sizes = set()
    with open(cache_path) as f:
        for line in f:
            parts = line.split()
            if len(parts) in (8, 9) and parts[0] == "gfx950":
                sizes.add(int(parts[1]))
    return sizes


DENSITY_NOISE, DENSITY_BAND = 0, 1


def init_fbm(n_theta, radius=1.0):
    """The reference's FBM curl-noise velocity (kam_init_fbm): FP32 (u_phi, u_theta)."""
    nt, nph = int(n_theta), 2 * int(n_theta)
    up = np.empty((nt, nph), np.float32)
    ut = np.empty((nt - 1, nph), np.float32)
    check(lib().kam_init_fbm(nt, nph, float(radius), _ptr(up), _ptr(ut)))
    return up, ut


def image_density(path, n_theta):
    """Density from a PPM/PGM image (kam_image_density): float32 [nTheta][nPhi]."""
    nt = int(n_theta)
    rho = np.empty((nt, 2 * nt), np.float32)
    check(lib().kam_image_density(str(path).encode(), nt, 2 * nt, _ptr(rho)))
    return rho


def image_colors(path, n_theta, coords):
    """Particle colours (G, R, B) / 255 from a PPM/PGM image (kam_image_colors)."""
    c = np.ascontiguousarray(coords, dtype=np.float32).reshape(-1, 2)
    out = np.empty((len(c), 3), np.float32)
    check(lib().kam_image_colors(str(path).encode(), int(n_theta), 2 * int(n_theta),
                                 _ptr(c) if len(c) else None, len(c), _ptr(out) if len(c) else None))
    return out


def particles_seed(n_theta, density, seed):
    """Jittered particle lattice (kam_particles_seed): float32 (n, 2) (phi, theta)."""
    n = C.c_int64(0)
    check(lib().kam_particles_seed(int(n_theta), float(density), int(seed), None, C.byref(n)))
    c = np.empty((n.value, 2), np.float32)
    check(lib().kam_particles_seed(int(n_theta), float(density), int(seed), _ptr(c), C.byref(n)))
    return c


def init_fields(n_theta, L, seed, density=DENSITY_BAND, radius=1.0):
    """Seeded synthetic state (kam_init_fields): FP32 (u_phi, u_theta, rho)."""
    nt, nph = int(n_theta), 2 * int(n_theta)
    up = np.empty((nt, nph), np.float32)
    ut = np.empty((nt - 1, nph), np.float32)
    rho = np.empty((nt, nph), np.float32)
    check(lib().kam_init_fields(nt, nph, float(radius), int(L), int(seed), int(density), _ptr(up),
                                _ptr(ut), _ptr(rho)))
    return up, ut, rho


class Sim:
    """One simulation on one GPU (the plain, one-sim-at-a-time form)."""

    def __init__(self, n_theta, dt, radius=1.0, geometric=True, device=0, precision=PREC_FP32):
        self.n_theta, self.n_phi = int(n_theta), 2 * int(n_theta)
        self.desc = SimDesc(self.n_theta, self.n_phi, float(radius), float(dt), int(bool(geometric)),
                            int(precision), int(device))
        h = C.c_void_p()
        check(lib().kam_sim_create(C.byref(self.desc), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().kam_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shapes(self):
        nt, nph = self.n_theta, self.n_phi
        return (nt, nph), (nt - 1, nph), (nt, nph)

    def set_fields(self, u_phi, u_theta, rho=None):
        sp, st, sr = self.shapes
        a, b = _f32(u_phi, sp), _f32(u_theta, st)
        c = None if rho is None else _f32(rho, sr)
        check(lib().kam_sim_set_fields(self._h, _ptr(a), _ptr(b), _ptr(c)))

    def get_fields(self):
        sp, st, sr = self.shapes
        a, b, c = np.empty(sp, np.float32), np.empty(st, np.float32), np.empty(sr, np.float32)
        check(lib().kam_sim_get_fields(self._h, _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c

    def step(self, nsteps=1):
        check(lib().kam_sim_step(self._h, int(nsteps)))

    def step_phases(self, phases):
        check(lib().kam_sim_step_phases(self._h, int(phases)))

    def spectrum(self):
        """Half spectrum, complex64 [M][nTheta] (mode-major; slot 0 = (mode 0, Nyquist))."""
        M = self.n_phi // 2
        s = np.empty((M, self.n_theta, 2), np.float32)
        check(lib().kam_sim_get_spectrum(self._h, _ptr(s)))
        return s[..., 0] + 1j * s[..., 1]

    def pressure(self):
        """Pressure of the last projection, float32 [nTheta][nPhi] (n = 0 mode removed)."""
        p = np.empty((self.n_theta, self.n_phi), np.float32)
        check(lib().kam_sim_get_pressure(self._h, _ptr(p)))
        return p

    def set_particles(self, coords):
        """Particles as (n, 2) (phi, theta) radians; later steps move them."""
        c = np.ascontiguousarray(coords, dtype=np.float32).reshape(-1, 2)
        check(lib().kam_sim_set_particles(self._h, _ptr(c) if len(c) else None, len(c)))

    def get_particles(self):
        n = lib().kam_sim_particle_count(self._h)
        c = np.empty((n, 2), np.float32)
        check(lib().kam_sim_get_particles(self._h, _ptr(c) if n else None, n))
        return c

    def poisson_residual(self):
        """Solves the current spectrum in place (as PHASE_TRI) and returns the
        per-mode relative residual of the solution, formed in FP64, float64 [M]."""
        r = np.empty(self.n_phi // 2, np.float64)
        mx, rms = C.c_double(), C.c_double()
        check(lib().kam_sim_poisson_residual(self._h, r.ctypes.data_as(C.POINTER(C.c_double)),
                                             C.byref(mx), C.byref(rms)))
        return r

    def sync(self):
        check(lib().kam_sim_sync(self._h))

    def status(self):
        """Watchdog word: 0 healthy, bit 0 once a step produced a non-finite value."""
        v = C.c_int(0)
        check(lib().kam_sim_status(self._h, C.byref(v)))
        return v.value

    def benchmark(self, nsteps, per_kernel=False):
        """Times nsteps steps with HIP events on the simulation's stream.
        Returns (total_ms, step_ms[nsteps], kernel_ms[4] or None)."""
        dp = C.POINTER(C.c_double)
        steps = np.zeros(nsteps, np.float64)
        kern = np.zeros(4, np.float64)
        tot = C.c_double(0.0)
        check(lib().kam_sim_benchmark(self._h, int(nsteps), int(bool(per_kernel)),
                                      steps.ctypes.data_as(dp), kern.ctypes.data_as(dp), C.byref(tot)))
        return tot.value, steps, (kern if per_kernel else None)


class KaminoSolver(Sim):
    """The reference's solver interface (KaminoSolver.cuh:99-111) over the C ABI.

    KaminoSolver(nPhi, nTheta, radius, frameDuration, ...) -> here the
    frame duration is not used by the step, and dt is fixed at construction as
    in the reference, whose kernels read the constant timeStepGlobal
    (KaminoCore.cu:871) and ignore stepForward's argument.
    """

    def __init__(self, n_phi, n_theta, radius, dt, geometric=True, device=0):
        if n_phi != 2 * n_theta:
            raise ValueError("nPhi must be 2 * nTheta")
        super().__init__(n_theta, dt, radius=radius, geometric=geometric, device=device)
        self.time_elapsed = 0.0

    def stepForward(self, time_step=None):  # noqa: N802 -- reference name
        self.step(1)
        self.time_elapsed += self.desc.dt if time_step is None else time_step


from .engine import Engine  # noqa: E402,F401
