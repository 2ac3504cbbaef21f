"""GPU test of kam_sim_step_noalloc, the step that allocates nothing: through
the C ABI, on device buffers the test owns, its results equal kam_sim_step's
(same kernels, same tables) bitwise, on calls in a row whose shapes shrink and
change, with and without density, in both precisions and on a stream of the
caller's; a shape beyond the scratch array is refused with a clear error and
leaves the next call correct."""
import ctypes as C

import numpy as np
import pytest

import kaminogpu_amd as K
from parity_util import log, seeded_state

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")

    def ok(self, rc, what):
        assert rc == 0, f"{what}: HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)), "hipMalloc")
        self.ok(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D), "hipMemcpy")
        return p

    def download(self, p, shape):
        out = np.empty(shape, np.float32)
        self.ok(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H), "hipMemcpy")
        return out

    def free(self, p):
        self.h.hipFree(p)


def solo(nt, dt, state, steps, precision):
    with K.Sim(nt, dt, precision=precision) as s:
        s.set_fields(*state)
        s.step(steps)
        return s.get_fields()


def noalloc(hip, nt, dt, state, steps, precision, stream=None, with_rho=True):
    up, ut, rho = state
    ptrs = [hip.upload(up), hip.upload(ut), hip.upload(rho) if with_rho else None]
    d = K.SimDesc(nt, 2 * nt, 1.0, float(dt), 1, int(precision), 0)
    try:
        K.check(K.lib().kam_sim_step_noalloc(C.byref(d), ptrs[0], ptrs[1], ptrs[2], int(steps), stream))
        hip.ok(hip.h.hipStreamSynchronize(stream), "hipStreamSynchronize")
        out = [hip.download(ptrs[0], up.shape), hip.download(ptrs[1], ut.shape)]
        out.append(hip.download(ptrs[2], rho.shape) if with_rho else None)
        return out
    finally:
        for p in ptrs:
            if p is not None:
                hip.free(p)


def test_noalloc_step_equals_sim_step():
    hip = Hip()
    stream = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        cases = [
            # (n_theta, steps, precision, stream, with rho): the second call is smaller
            (256, 3, K.PREC_FP32, None, True),
            (48, 2, K.PREC_FP32, None, True),
            (75, 2, K.PREC_FP64_PRESSURE, stream, True),
            (64, 1, K.PREC_FP32, stream, False),
            (64, 2, K.PREC_FP32, None, True),  # same shape again: no re-upload
        ]
        for nt, steps, prec, st, with_rho in cases:
            dt = 0.5 / nt
            state = seeded_state(nt, L=min(16, nt // 4), seed=40 + nt, dens=1)
            log(f"[noalloc] n_theta={nt} steps={steps} precision={prec} "
                f"stream={'own' if st else 'null'} rho={with_rho}")
            got = noalloc(hip, nt, dt, state, steps, prec, st, with_rho)
            ref_state = state if with_rho else (state[0], state[1], np.zeros_like(state[2]))
            ref = solo(nt, dt, ref_state, steps, prec)
            np.testing.assert_array_equal(got[0], ref[0])
            np.testing.assert_array_equal(got[1], ref[1])
            if with_rho:
                np.testing.assert_array_equal(got[2], ref[2])
        # beyond the scratch array: refused, and the next call is still right
        big = K.SimDesc(1024, 2048, 1.0, 0.5 / 1024, 1, 0, 0)
        rc = K.lib().kam_sim_step_noalloc(C.byref(big), 1, 1, None, 1, None)
        msg = K.lib().kam_last_error().decode()
        log(f"[noalloc] n_theta=1024 -> {rc}: {msg}")
        assert rc == -1 and "scratch" in msg
        nt = 32
        state = seeded_state(nt, L=8, seed=77, dens=1)
        got = noalloc(hip, nt, 0.5 / nt, state, 2, K.PREC_FP32)
        for a, b in zip(got, solo(nt, 0.5 / nt, state, 2, K.PREC_FP32)):
            np.testing.assert_array_equal(a, b)
    finally:
        hip.h.hipStreamDestroy(stream)
