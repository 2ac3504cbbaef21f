"""Particles: the oracle's particle step against closed forms (CPU), the seeded
layout (CPU), and the HIP particle kernel against the oracle (GPU).

advectionParticles (KaminoCore.cu:321-342) moves each particle one forward-
Euler step with the pre-step velocities sampled at the particle, then folds
the position with validateCoord.  Positions are compared on the sphere
(Cartesian, radius 1), which is continuous across the pole fold and the phi
seam; the tolerance is the per-step 1e-5 relative L2 of the fields.
"""
import math

import numpy as np
import pytest

import kaminogpu_amd as K
from oracle import oracle as O

TOL_POS = 1e-5


def cart(pos):
    p = np.asarray(pos, np.float64)
    phi, th = p[:, 0], p[:, 1]
    return np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], 1)


def rel_pos(a, b, mask=None):
    d = cart(a) - cart(b)
    keep = np.ones(len(d), bool) if mask is None else ~mask.astype(bool)
    return float(np.linalg.norm(d[keep]) / np.linalg.norm(cart(b)[keep]))


def rotation_field(nt, omega, radius=1.0):
    """u_phi = omega R sin(theta) at the u_phi nodes (theta = (j + 1/2) h), u_theta = 0."""
    h = math.pi / nt
    th = (np.arange(nt) + 0.5) * h
    up = np.repeat((omega * radius * np.sin(th))[:, None], 2 * nt, 1)
    return up, np.zeros((nt - 1, 2 * nt))


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------
def test_oracle_zero_velocity_identity():
    nt = 16
    rng = np.random.default_rng(0)
    pos = np.stack([rng.uniform(0, 2 * math.pi, 200), rng.uniform(0.01, math.pi - 0.01, 200)], 1)
    out, _ = O.advect_particles(nt, 2 * nt, 0.01, np.zeros((nt, 2 * nt)), np.zeros((nt - 1, 2 * nt)), pos)
    np.testing.assert_allclose(out, pos, rtol=0, atol=1e-15)


def test_oracle_rotation_on_rows_is_exact():
    """On a u_phi row the sample is the row value omega R sin(theta), so phi
    advances by omega dt exactly and theta is unchanged."""
    nt, omega, dt = 32, 0.7, 0.01
    h = math.pi / nt
    up, ut = rotation_field(nt, omega)
    j = np.arange(2, nt - 2)
    phi = np.linspace(0.1, 6.0, len(j))
    pos = np.stack([phi, (j + 0.5) * h], 1)
    out, _ = O.advect_particles(nt, 2 * nt, dt, up, ut, pos)
    np.testing.assert_allclose(out[:, 0], phi + omega * dt, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[:, 1], pos[:, 1], rtol=0, atol=1e-15)


def test_oracle_rotation_between_rows_interpolates_linearly():
    """Between rows j and j+1 the u_phi sample is the linear interpolation in
    theta of the two rows: phi advances by omega dt lerp(sin t_j, sin t_j+1, a) / sin theta."""
    nt, omega, dt = 32, 0.7, 0.01
    h = math.pi / nt
    up, ut = rotation_field(nt, omega)
    a = np.array([0.1, 0.25, 0.5, 0.8])
    j = np.array([3, 9, 17, 26])
    th = (j + 0.5 + a) * h
    pos = np.stack([np.full(4, 1.0), th], 1)
    out, _ = O.advect_particles(nt, 2 * nt, dt, up, ut, pos)
    lerp = (1 - a) * np.sin((j + 0.5) * h) + a * np.sin((j + 1.5) * h)
    np.testing.assert_allclose(out[:, 0], 1.0 + omega * dt * lerp / np.sin(th), rtol=0, atol=1e-12)


def test_oracle_constant_meridional_flow_and_pole_fold():
    """u_theta = c everywhere: theta moves by c dt / R; a particle pushed past
    the north pole comes back at -theta on the antipodal meridian."""
    nt, c, dt = 16, -2.0, 0.01
    ut = np.full((nt - 1, 2 * nt), c)
    up = np.zeros((nt, 2 * nt))
    pos = np.array([[0.5, 1.0], [2.0, 0.005], [4.0, 0.015]])
    out, _ = O.advect_particles(nt, 2 * nt, dt, up, ut, pos)
    assert out[0, 1] == pytest.approx(1.0 + c * dt, abs=1e-15)
    assert out[0, 0] == pytest.approx(0.5, abs=1e-15)
    # 0.005 - 0.02 = -0.015 -> 0.015, phi + pi
    assert out[1, 1] == pytest.approx(0.015, abs=1e-15)
    assert out[1, 0] == pytest.approx(2.0 + math.pi, abs=1e-12)
    assert out[2, 1] == pytest.approx(0.005, abs=1e-15)
    assert out[2, 0] == pytest.approx(4.0 + math.pi - 2 * math.pi, abs=1e-12)


def test_oracle_no_phi_motion_at_the_pole():
    """R sin(theta) <= 1e-7 leaves phi alone whatever u_phi is."""
    nt = 16
    up = np.full((nt, 2 * nt), 3.0)
    out, _ = O.advect_particles(nt, 2 * nt, 0.01, up, np.zeros((nt - 1, 2 * nt)),
                                np.array([[1.25, 0.0]]))
    assert out[0, 0] == 1.25 and out[0, 1] == 0.0


# ---------------------------------------------------------------------------
# seeded layout (CPU, host code of the library)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nt,density", [(16, 4.0), (32, 1.0), (24, 9.0)])
def test_particles_seed_layout(nt, density):
    c = K.particles_seed(nt, density, 11)
    lin = math.sqrt(density)
    n_th = int(np.float32(lin) * np.float32(nt))
    assert c.shape == (2 * n_th * n_th, 2)
    delta = math.pi / nt / lin
    i, j = np.divmod(np.arange(len(c)), n_th)
    # each particle within half a lattice step of its lattice point (or clamped at 0)
    assert np.all(np.abs(c[:, 0] - i * delta) <= delta / 2 * (1 + 1e-5) + 1e-6)
    assert np.all(np.abs(c[:, 1] - j * delta) <= delta / 2 * (1 + 1e-5) + 1e-6)
    assert c.min() >= 0.0
    np.testing.assert_array_equal(c, K.particles_seed(nt, density, 11))
    assert not np.array_equal(c, K.particles_seed(nt, density, 12))


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def _gpu_particles(nt, dt, state, pos, steps=1):
    with K.Sim(nt, dt) as s:
        s.set_fields(*state)
        s.set_particles(pos)
        out = []
        for _ in range(steps):
            f = s.get_fields()
            s.step(1)
            out.append((f, s.get_particles()))
        return out


@pytest.mark.gpu
@pytest.mark.parametrize("nt", [32, 64, 256])
def test_gpu_particles_vs_oracle(nt):
    dt = 0.5 / nt
    up, ut, rho = [np.ascontiguousarray(a, np.float32) for a in O.init_fields(nt, 2 * nt, min(16, nt // 4), nt, 1)]
    pos = K.particles_seed(nt, 2.0, 3)
    print(f"[particles] nt={nt} n={len(pos)}", flush=True)
    res = _gpu_particles(nt, dt, (up, ut, rho), pos, steps=5)
    prev = pos
    for k, ((f_up, f_ut, _), got) in enumerate(res):
        ref, amb = O.advect_particles(nt, 2 * nt, dt, f_up, f_ut, prev)
        err = rel_pos(got, ref, amb)
        # raw coordinates too, away from the fold and the seam
        ok = (~amb.astype(bool)) & (np.abs(ref[:, 0]) > 1e-4) & (np.abs(ref[:, 0] - 2 * math.pi) > 1e-4) \
            & (ref[:, 1] > 1e-4) & (ref[:, 1] < math.pi - 1e-4)
        raw = float(np.linalg.norm(got[ok] - ref[ok]) / np.linalg.norm(ref[ok]))
        print(f"  step {k}: rel L2 (sphere) {err:.2e}, raw {raw:.2e}, ambiguous {int(amb.sum())}", flush=True)
        assert err < TOL_POS and raw < TOL_POS
        prev = got


@pytest.mark.gpu
def test_gpu_particles_rotation_and_fold():
    nt, omega, dt = 64, 0.7, 0.005
    h = math.pi / nt
    up, ut = rotation_field(nt, omega)
    rho = np.zeros((nt, 2 * nt))
    j = np.arange(2, nt - 2)
    pos = np.stack([np.linspace(0.1, 6.0, len(j)), (j + 0.5) * h], 1).astype(np.float32)
    with K.Sim(nt, dt, geometric=False) as s:
        s.set_fields(*[np.ascontiguousarray(a, np.float32) for a in (up, ut, rho)])
        s.set_particles(pos)
        s.step_phases(K.PHASE_ADVECT)
        got = s.get_particles()
    np.testing.assert_allclose(got[:, 0], pos[:, 0] + omega * dt, rtol=0, atol=2e-6)
    np.testing.assert_array_equal(got[:, 1], pos[:, 1])
    # meridional flow through the north pole
    ut2 = np.full((nt - 1, 2 * nt), -2.0, np.float32)
    p2 = np.array([[2.0, 0.005], [0.5, 1.0]], np.float32)
    with K.Sim(nt, 0.01, geometric=False) as s:
        s.set_fields(np.zeros((nt, 2 * nt), np.float32), ut2, np.zeros((nt, 2 * nt), np.float32))
        s.set_particles(p2)
        s.step_phases(K.PHASE_ADVECT)
        got2 = s.get_particles()
    assert got2[0, 1] == pytest.approx(0.015, abs=5e-7)  # the fold rounds theta + 2 pi to FP32, as the reference
    assert got2[0, 0] == pytest.approx(2.0 + math.pi, abs=1e-6)
    assert got2[1, 1] == pytest.approx(0.98, abs=2e-7)


@pytest.mark.gpu
def test_gpu_particles_leave_fields_alone():
    nt, dt = 64, 0.5 / 64
    state = [np.ascontiguousarray(a, np.float32) for a in O.init_fields(nt, 2 * nt, 8, 5, 1)]
    with K.Sim(nt, dt) as s:
        s.set_fields(*state)
        s.step(3)
        plain = s.get_fields()
    with K.Sim(nt, dt) as s:
        s.set_fields(*state)
        s.set_particles(K.particles_seed(nt, 4.0, 1))
        s.step(3)
        with_p = s.get_fields()
        s.set_particles(np.zeros((0, 2), np.float32))
        assert len(s.get_particles()) == 0
        s.step(1)
    for a, b in zip(plain, with_p):
        np.testing.assert_array_equal(a, b)
