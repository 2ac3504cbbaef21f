"""Shared helpers of the GPU parity tests: seeded states, oracle references and
the error metrics (relative L2 over cells the oracle did not flag as
branch-ambiguous)."""
import sys

import numpy as np

from oracle import oracle as O

TOL_FIELD = 1e-5  # per-step relative L2 of advected / projected fields


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def f32(*arrs):
    return [np.ascontiguousarray(a, dtype=np.float32) for a in arrs]


def f64(*arrs):
    return [np.asarray(a, dtype=np.float64) for a in arrs]


def seeded_state(nt, L=8, seed=1, dens=1, pre_steps=0, dt=None, geometric=True):
    """A seeded state rounded to FP32 (the exact input both sides start from).
    pre_steps oracle steps make the velocity generic (no longer a pure curl)."""
    nph = 2 * nt
    up, ut, rho = O.init_fields(nt, nph, L, seed, dens)
    if pre_steps:
        dt = dt if dt is not None else 0.5 / nt
        ph = O.PHASE_ALL if geometric else (O.PHASE_ADVECT | O.PHASE_PROJECT)
        for _ in range(pre_steps):
            up, ut, rho, _ = O.step(nt, nph, dt, up, ut, rho, phases=ph)
    return f32(up, ut, rho)


def oracle_advect_geo(nt, dt, up, ut, rho, geometric=True, tol=1e-5):
    nph = 2 * nt
    a_up, a_ut, a_rho, m_up, m_ut, m_rho, st = O.advect(nt, nph, dt, *f64(up, ut, rho), tol=tol)
    if geometric:
        a_up, a_ut = O.geometric(nt, nph, dt, a_up, a_ut)
        m_up, m_ut = O.geometric_mask(nt, nph, m_up, m_ut)
    return (a_up, a_ut, a_rho), (m_up, m_ut, m_rho), st


def rel_l2(a, b, mask=None):
    return O.rel_l2(a, b, mask)


def div_metric(nt, up_gpu, ut_gpu, up_ref, ut_ref, up_pre, ut_pre):
    """||D(u_gpu) - D(u_ref)|| / ||D(u_pre)|| with D the reference divergence."""
    nph = 2 * nt
    d_gpu = O.divergence(nt, nph, *f64(up_gpu, ut_gpu))
    d_ref = O.divergence(nt, nph, *f64(up_ref, ut_ref))
    d_pre = O.divergence(nt, nph, *f64(up_pre, ut_pre))
    return float(np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_pre))
