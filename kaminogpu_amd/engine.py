"""Python view of the batching / serving engine (kam_engine_* in kamino.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

ENGINE_SYMBOLS = [
    "kam_engine_create", "kam_engine_destroy", "kam_engine_admit", "kam_engine_step",
    "kam_engine_poll", "kam_engine_fetch", "kam_engine_cancel", "kam_engine_sync",
    "kam_engine_stats", "kam_engine_reset_stats", "kam_engine_profile", "kam_engine_now",
    "kam_engine_completed",
]

REQ_ACTIVE, REQ_DONE, REQ_CANCELLED, REQ_FAILED = 1, 2, 3, 4
KAM_EAGAIN = -4


class EngineDesc(C.Structure):
    _fields_ = [("device", C.c_int), ("max_slots", C.c_int), ("hbm_budget_bytes", C.c_size_t),
                ("use_graph", C.c_int), ("max_inflight", C.c_int), ("graph_steps", C.c_int),
                ("groups", C.c_int)]


class ReqStatus(C.Structure):
    _fields_ = [("state", C.c_int), ("reserved", C.c_int), ("steps_done", C.c_int64),
                ("steps_total", C.c_int64), ("admit_s", C.c_double), ("done_s", C.c_double)]


class EngineStats(C.Structure):
    _fields_ = [("active", C.c_int), ("reserved", C.c_int), ("steps", C.c_int64),
                ("cell_steps", C.c_int64), ("gpu_ms", C.c_double), ("step_p50_ms", C.c_double),
                ("step_p99_ms", C.c_double), ("step_max_ms", C.c_double), ("admitted", C.c_int64),
                ("completed", C.c_int64), ("rejected", C.c_int64), ("cancelled", C.c_int64),
                ("bytes_in_use", C.c_size_t), ("bytes_budget", C.c_size_t), ("failed", C.c_int64)]


def bind(L):
    from . import SimDesc

    fp = C.POINTER(C.c_float)
    vp = C.c_void_p
    L.kam_engine_create.argtypes = [C.POINTER(EngineDesc), C.POINTER(vp)]
    L.kam_engine_destroy.argtypes = [vp]
    L.kam_engine_admit.argtypes = [vp, C.POINTER(SimDesc), fp, fp, fp, C.c_int64, C.POINTER(C.c_uint64)]
    L.kam_engine_step.argtypes = [vp, C.c_int]
    L.kam_engine_poll.argtypes = [vp, C.c_uint64, C.POINTER(ReqStatus)]
    L.kam_engine_fetch.argtypes = [vp, C.c_uint64, fp, fp, fp]
    L.kam_engine_cancel.argtypes = [vp, C.c_uint64]
    L.kam_engine_sync.argtypes = [vp]
    L.kam_engine_stats.argtypes = [vp, C.POINTER(EngineStats)]
    L.kam_engine_reset_stats.argtypes = [vp]
    L.kam_engine_profile.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.kam_engine_now.restype = C.c_double
    L.kam_engine_now.argtypes = [vp]
    L.kam_engine_completed.argtypes = [vp, C.POINTER(C.c_uint64), C.c_int, C.POINTER(C.c_int)]


class Engine:
    """Batched / served stepping of many simulations on one GPU.  Steps go out
    as direct kernel launches by default; use_graph=True replays them from
    hipGraphs instead (same results).  Each graph launch boundary costs the GPU
    ~10 us, so graph_steps=G > 1 captures G steps per graph and replays it
    whenever no simulation finishes within the next G steps (the steps between
    such runs are launched directly).  groups=G > 1 splits the simulations into
    G groups stepped on streams of their own, whose kernels overlap."""

    def __init__(self, device=0, max_slots=0, hbm_budget_bytes=0, use_graph=False, max_inflight=0,
                 graph_steps=1, groups=1):
        from . import check, lib

        self._lib = lib()
        self._check = check
        d = EngineDesc(int(device), int(max_slots), int(hbm_budget_bytes), int(bool(use_graph)),
                       int(max_inflight), int(graph_steps), int(groups))
        h = C.c_void_p()
        check(self._lib.kam_engine_create(C.byref(d), C.byref(h)))
        self._h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kam_engine_destroy(self._h)
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

    def admit(self, n_theta, dt, u_phi, u_theta, rho, steps, radius=1.0, geometric=True, precision=0):
        """Returns the request id, or None when the engine has no room now.
        precision: PREC_FP32 or PREC_FP64_PRESSURE."""
        from . import SimDesc, _f32, _ptr

        nt, nph = int(n_theta), 2 * int(n_theta)
        sd = SimDesc(nt, nph, float(radius), float(dt), int(bool(geometric)), int(precision), self.device)
        a, b = _f32(u_phi, (nt, nph)), _f32(u_theta, (nt - 1, nph))
        c = None if rho is None else _f32(rho, (nt, nph))
        rid = C.c_uint64(0)
        rc = self._lib.kam_engine_admit(self._h, C.byref(sd), _ptr(a), _ptr(b), _ptr(c),
                                        int(steps), C.byref(rid))
        if rc == KAM_EAGAIN:
            return None
        self._check(rc)
        return rid.value

    def step(self, nsteps=1):
        self._check(self._lib.kam_engine_step(self._h, int(nsteps)))

    def poll(self, rid):
        st = ReqStatus()
        self._check(self._lib.kam_engine_poll(self._h, C.c_uint64(rid), C.byref(st)))
        return st

    def fetch(self, rid, n_theta):
        from . import _ptr

        nt, nph = int(n_theta), 2 * int(n_theta)
        a = np.empty((nt, nph), np.float32)
        b = np.empty((nt - 1, nph), np.float32)
        c = np.empty((nt, nph), np.float32)
        self._check(self._lib.kam_engine_fetch(self._h, C.c_uint64(rid), _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c

    def fetch_release(self, rid):
        """Forgets a finished request without copying its fields."""
        self._check(self._lib.kam_engine_fetch(self._h, C.c_uint64(rid), None, None, None))

    def cancel(self, rid):
        self._check(self._lib.kam_engine_cancel(self._h, C.c_uint64(rid)))

    def sync(self):
        self._check(self._lib.kam_engine_sync(self._h))

    def stats(self):
        s = EngineStats()
        self._check(self._lib.kam_engine_stats(self._h, C.byref(s)))
        return s

    def reset_stats(self):
        self._check(self._lib.kam_engine_reset_stats(self._h))

    def profile(self, nsteps):
        k = (C.c_double * 4)()
        self._check(self._lib.kam_engine_profile(self._h, int(nsteps), k))
        return np.array(list(k))

    def now(self):
        return self._lib.kam_engine_now(self._h)

    def completed(self, max_ids=4096):
        """Ids whose final fields reached the host since the last call."""
        buf = (C.c_uint64 * max_ids)()
        n = C.c_int(0)
        self._check(self._lib.kam_engine_completed(self._h, buf, int(max_ids), C.byref(n)))
        return [buf[k] for k in range(n.value)]
