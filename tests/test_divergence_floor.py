"""CPU evidence for the divergence gate (ii) bar, with the oracle: the FP32
storage floor ||D(fl32(u_post)) - D(u_post)|| / ||D(u_pre)|| is small, an
FP32 pressure pipeline (d-hat, p-hat and the inverse FFT in FP32) sits far
above it, and keeping p-hat, the inverse FFT and the gradient in FP64 (the
FP64-pressure mode's arithmetic) lands back on the floor.  Pipelines are
emulated with numpy FFTs (complex64 for the FP32 stages) and dense per-mode
solves of the oracle's Poisson rows."""
import numpy as np
import pytest

from oracle import oracle as O
from parity_util import f32, f64, oracle_advect_geo, seeded_state


def D(nt, up, ut):
    return O.divergence(nt, 2 * nt, *f64(up, ut))


def grad_sub(nt, up, ut, p):
    h = np.pi / nt
    th = (np.arange(nt) + 0.5) * h
    up = up - (p - np.roll(p, 1, axis=1)) / (h * np.sin(th))[:, None]
    ut = ut - (p[1:] - p[:-1]) / h
    return up, ut


def emulate(nt, up, ut, fft32):
    """d-hat always from an FP32 FFT; p-hat, the inverse FFT and the gradient
    in FP32 (fft32) or FP64."""
    nph, M = 2 * nt, nt
    d = D(nt, up, ut)
    dh = (np.fft.rfft(d.astype(np.float32), axis=1) / np.float32(nph)).astype(np.complex64).astype(np.complex128)
    ph = np.zeros_like(dh)
    for n in range(1, M + 1):
        a, b, c = O.poisson_rows(nt, nph, n)
        T = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        ph[:, n] = np.linalg.solve(T, dh[:, n])
    if fft32:
        p = np.fft.irfft(ph.astype(np.complex64) * np.float32(nph), n=nph, axis=1).astype(np.float64)
    else:
        p = np.fft.irfft(ph * nph, n=nph, axis=1)
    return f32(*grad_sub(nt, f64(up)[0], f64(ut)[0], p))


@pytest.mark.parametrize("nt", [32, 64])
def test_fp64_pressure_reaches_the_storage_floor(nt):
    dt = 0.5 / nt
    state = seeded_state(nt, L=max(1, min(16, nt // 4)), seed=3 * nt, dens=1)
    (up, ut, _), _, _ = oracle_advect_geo(nt, dt, *state)
    up, ut = f32(up, ut)
    r_up, r_ut, _, _ = O.project(nt, 2 * nt, *f64(up, ut))
    dpre = np.linalg.norm(D(nt, up, ut))
    dref = D(nt, r_up, r_ut)

    def gate(a, b):
        return np.linalg.norm(D(nt, a, b) - dref) / dpre

    floor = gate(*f32(r_up, r_ut))
    g32 = gate(*emulate(nt, up, ut, True))
    g64 = gate(*emulate(nt, up, ut, False))
    print(f"nt={nt} storage floor {floor:.2e}  FP32 pressure {g32:.2e}  FP64 pressure {g64:.2e}")
    assert floor < 1e-6
    assert g32 > 10 * floor
    assert g64 < 1.5 * floor
