"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Each check comes from an independent route: numpy/scipy goldens (committed
under tests/golden/), published known-answer vectors, closed forms of the
mathematics, invariants, and sampler values derived by hand from the reference
source (KaminoCore.cu:11-184).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O


# ---------------------------------------------------------------- PRNG ----
def test_splitmix64_published_vectors():
    # Published SplitMix64 outputs (seed 0, and the five values for seed 1234567).
    s = C.c_uint64(0)
    assert O.lib().ko_splitmix64(C.byref(s)) == 0xE220A8397B1DCDAF
    s = C.c_uint64(1234567)
    got = [O.lib().ko_splitmix64(C.byref(s)) for _ in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_normal_moments():
    s = C.c_uint64(7)
    x = np.array([O.lib().ko_normal(C.byref(s)) for _ in range(40000)])
    assert abs(x.mean()) < 0.02 and abs(x.std() - 1.0) < 0.02


# ------------------------------------------------------------ harmonics ----
@pytest.mark.parametrize("l,m", [(1, 0), (1, 1), (2, -1), (4, 3), (7, -5), (16, 16), (32, -31)])
def test_real_sh_against_scipy(l, m):
    rng = np.random.default_rng(l * 100 + m)
    for th, ph in rng.uniform([0.01, 0.0], [np.pi - 0.01, 2 * np.pi], size=(5, 2)):
        Y = sp.sph_harm_y(l, abs(m), th, ph)  # includes the Condon-Shortley phase
        if m > 0:
            ref = np.sqrt(2) * (-1) ** m * Y.real
        elif m < 0:
            ref = np.sqrt(2) * (-1) ** abs(m) * Y.imag
        else:
            ref = Y.real
        assert O.lib().ko_real_sh(l, m, th, ph) == pytest.approx(ref, rel=1e-10, abs=1e-12)


def test_init_velocity_is_discretely_divergence_free():
    for nt, L, seed in [(16, 4, 3), (64, 8, 1), (48, 12, 9)]:
        up, ut, rho = O.init_fields(nt, 2 * nt, L, seed, 1)
        d = O.divergence(nt, 2 * nt, up, ut)
        scale = np.abs(up).max() / (np.pi / nt)
        assert np.abs(d).max() < 1e-12 * scale
        assert np.abs(up).max() > 1e-3


def test_init_density_kinds():
    nt = 64
    h = np.pi / nt
    _, _, band = O.init_fields(nt, 2 * nt, 8, 5, 1)
    th = (np.arange(nt) + 0.5) * h
    expect = np.clip(0.5 + (0.3 - np.abs(th - np.pi / 2)) / (2 * h), 0, 1)
    np.testing.assert_allclose(band, np.repeat(expect[:, None], 2 * nt, 1), rtol=0, atol=1e-15)
    _, _, noise = O.init_fields(nt, 2 * nt, 8, 5, 0)
    assert noise.min() >= 0.5 and noise.max() < 0.6 and noise.std() > 0.02


# ------------------------------------------------------------------ FFT ----
def test_fft_against_numpy_goldens(goldens):
    for n in goldens["dft_lengths"]:
        x = goldens[f"dft_in_{n}"]
        for sign, key in [(-1, "fwd"), (1, "inv")]:
            ref = goldens[f"dft_{key}_{n}"]
            got = O.fft(x, sign)
            assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (n, sign)


def test_fft_closed_form_single_mode():
    n, k = 48, 5
    i = np.arange(n)
    X = O.fft(np.cos(2 * np.pi * k * i / n), -1)
    expect = np.zeros(n, complex)
    expect[k] = expect[n - k] = n / 2
    np.testing.assert_allclose(X, expect, atol=1e-12)


# ---------------------------------------------------------- tridiagonal ----
def test_poisson_rows_and_solves_match_goldens(goldens):
    for nt, n in goldens["tri_cases"]:
        a, b, c = O.poisson_rows(int(nt), 2 * int(nt), int(n))
        np.testing.assert_allclose(a, goldens[f"tri_a_{nt}_{n}"], rtol=1e-14, atol=1e-12)
        np.testing.assert_allclose(b, goldens[f"tri_b_{nt}_{n}"], rtol=1e-14, atol=1e-12)
        np.testing.assert_allclose(c, goldens[f"tri_c_{nt}_{n}"], rtol=1e-14, atol=1e-12)
        d = goldens[f"tri_d_{nt}_{n}"]
        x = goldens[f"tri_x_{nt}_{n}"]
        np.testing.assert_allclose(O.thomas(a, b, c, d), x, rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(O.cyclic_reduction(a, b, c, d), x, rtol=1e-11, atol=1e-13)


def test_cyclic_reduction_equals_thomas_large():
    rng = np.random.default_rng(4)
    for nt in (64, 256, 1024):
        for n in (1, 7, nt):
            a, b, c = O.poisson_rows(nt, 2 * nt, n)
            d = rng.standard_normal(nt)
            xt = O.thomas(a, b, c, d)
            xc = O.cyclic_reduction(a, b, c, d)
            assert np.abs(xt - xc).max() <= 1e-11 * np.abs(xt).max()
    with pytest.raises(ValueError):
        a, b, c = O.poisson_rows(48, 96, 3)
        O.cyclic_reduction(a, b, c, np.ones(48))


@pytest.mark.parametrize("l,m", [(2, 1), (3, 1), (6, 4)])  # mode 0 is the identity row
def test_poisson_rows_are_second_order_sphere_laplacian(l, m):
    # Interior rows applied to P_l^m(cos theta) give -l(l+1) P_l^m + O(h^2):
    # halving h cuts the error by ~4 (the a/b/c signs and the cot term are pinned).
    errs = []
    for nt in (64, 128):
        h = np.pi / nt
        th = (np.arange(nt) + 0.5) * h
        f = sp.lpmv(m, l, np.cos(th))
        a, b, c = O.poisson_rows(nt, 2 * nt, m)
        Tf = b * f
        Tf[1:] += a[1:] * f[:-1]
        Tf[:-1] += c[:-1] * f[1:]
        inner = (th > 0.3) & (th < np.pi - 0.3)  # fixed in theta: cot is O(1) there
        errs.append(np.abs(Tf[inner] + l * (l + 1) * f[inner]).max() / np.abs(f).max())
    assert errs[1] < errs[0] / 3.0 and errs[1] < 1e-2 * l * (l + 1)


# ----------------------------------------------------------- projection ----
def test_projection_matches_dense_goldens(goldens):
    for nt in goldens["proj_sizes"]:
        nt = int(nt)
        nph = 2 * nt
        up_in = goldens[f"proj_uphi_in_{nt}"]
        ut_in = goldens[f"proj_utheta_in_{nt}"]
        np.testing.assert_allclose(O.divergence(nt, nph, up_in, ut_in), goldens[f"proj_div_{nt}"],
                                   rtol=1e-13, atol=1e-12)
        for solver in ((0, 1) if nt & (nt - 1) == 0 else (0,)):
            up, ut, spec, p = O.project(nt, nph, up_in, ut_in, solver=solver)
            np.testing.assert_allclose(spec.T, goldens[f"proj_dhat_{nt}"], rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(p, goldens[f"proj_p_{nt}"], rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(up, goldens[f"proj_uphi_out_{nt}"], rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(ut, goldens[f"proj_utheta_out_{nt}"], rtol=1e-10, atol=1e-12)


def test_projection_is_identity_on_curl_fields():
    nt = 32
    up, ut, _ = O.init_fields(nt, 2 * nt, 8, 11, 0)
    up2, ut2, _, p = O.project(nt, 2 * nt, up, ut)
    assert np.abs(p).max() < 1e-12
    np.testing.assert_allclose(up2, up, atol=1e-13)
    np.testing.assert_allclose(ut2, ut, atol=1e-13)


def test_projection_drops_zero_mode_and_is_real_linear():
    nt = 16
    rng = np.random.default_rng(2)
    u1 = rng.standard_normal((nt, 2 * nt)), rng.standard_normal((nt - 1, 2 * nt))
    u2 = rng.standard_normal((nt, 2 * nt)), rng.standard_normal((nt - 1, 2 * nt))
    _, _, _, p1 = O.project(nt, 2 * nt, *u1)
    _, _, _, p2 = O.project(nt, 2 * nt, *u2)
    _, _, _, p12 = O.project(nt, 2 * nt, 2 * u1[0] - u2[0], 2 * u1[1] - u2[1])
    np.testing.assert_allclose(p12, 2 * p1 - p2, atol=1e-10)
    assert np.abs(p1.mean(axis=1)).max() < 1e-12  # the n = 0 mode is removed


# ---------------------------------------------------------------- cubic ----
def test_cubic_against_numpy_roots(goldens):
    for G, v, u, x in zip(goldens["cubic_G"], goldens["cubic_v"], goldens["cubic_u"], goldens["cubic_x"]):
        cof = G * G
        got = O.solve_cubic(0.0, (G * v + 1.0) / cof, -u / cof)
        assert got == pytest.approx(x, rel=1e-8, abs=1e-10)
        assert abs(G * G * got**3 + (1 + G * v) * got - u) < 1e-9 * max(1, abs(u))


def test_cubic_trig_branch_returns_a_root():
    # 1 + G v < 0 gives three real roots; the reference takes its trig root.
    G, v, u = 0.5, -6.0, 0.3
    B, Cc = (G * v + 1.0) / (G * G), -u / (G * G)
    x = O.solve_cubic(0.0, B, Cc)
    assert abs(x**3 + B * x + Cc) < 1e-10 * max(1, abs(B))


def test_cubic_special_cases():
    assert abs(O.solve_cubic(0.0, 1e6, 0.0)) < 1e-9  # u = 0 -> x = 0 (Cardano cancels to ~1e-13)
    G = 1e-4
    x = O.solve_cubic(0.0, (G * 0.2 + 1) / G**2, -0.7 / G**2)
    assert x == pytest.approx(0.7 / (1 + G * 0.2), rel=1e-6)


def test_cubic_overflowed_input_returns_non_finite():
    # an infinite u_phi in the one-real-root Cardano branch: the cube root's
    # scaling loop would never end on +inf; the kernel carries the same guard
    for c in (np.inf, -np.inf):
        assert not np.isfinite(O.solve_cubic(0.0, -4.0, c))


# ------------------------------------------------------------ geometric ----
def test_geometric_closed_form_at_dt_zero():
    nt, nph = 8, 16
    rng = np.random.default_rng(3)
    up = rng.standard_normal((nt, nph))
    ut = rng.standard_normal((nt - 1, nph))
    gp, gt = O.geometric(nt, nph, 0.0, up, ut)
    np.testing.assert_allclose(gp, 0.25 * (np.roll(up, 1, 1) + 2 * up + np.roll(up, -1, 1)), atol=1e-14)
    vbar = np.zeros((nt, nph))
    vbar[1:-1] = 0.5 * (ut[:-1] + ut[1:])
    vbar[0] = 0.75 * ut[0] + 0.25 * np.roll(ut[0], -nph // 2)
    vbar[-1] = 0.75 * ut[-1] + 0.25 * np.roll(ut[-1], -nph // 2)
    np.testing.assert_allclose(gt, 0.5 * (vbar[:-1] + vbar[1:]), atol=1e-14)


def test_geometric_zonal_flow_at_dt():
    """Zonal flow u_phi = f(theta_j), u_theta = 0, at dt > 0 and R != 1: every
    cell of row j has ubar = f_j and vbar = 0 (also in the pole rows' 3/4, 1/4
    antipodal mix), so x_j is the one real root of G_j^2 x^3 + x - f_j with
    G_j = dt cot(theta_j) / R (KaminoCore.cu:494-510), taken here from
    numpy.roots; u_phi' = x_j (assignPhiKernel :517-534) and
    u_theta'_j = (G_j x_j^2 + G_{j+1} x_{j+1}^2) / 2 (v' = vbar + G x^2,
    assignThetaKernel :537-549)."""
    nt, nph, dt, R = 16, 32, 0.05, 2.0
    h = np.pi / nt
    th = (np.arange(nt) + 0.5) * h
    f = 1.0 + 0.5 * np.cos(th) + 0.3 * np.sin(3 * th)
    up = np.repeat(f[:, None], nph, 1)
    ut = np.zeros((nt - 1, nph))
    gp, gt = O.geometric(nt, nph, dt, up, ut, radius=R)
    G = dt * np.cos(th) / (R * np.sin(th))
    x = np.empty(nt)
    for j in range(nt):
        if abs(G[j]) <= 1e-7:
            x[j] = f[j]
            continue
        r = np.roots([G[j] ** 2, 0.0, 1.0, -f[j]])
        x[j] = r[np.argmin(np.abs(r.imag))].real
    assert np.abs(G).max() > 0.2  # the cubic term matters near the poles
    np.testing.assert_allclose(gp, np.repeat(x[:, None], nph, 1), rtol=1e-10, atol=1e-12)
    vp = G * x ** 2
    np.testing.assert_allclose(gt, np.repeat((0.5 * (vp[:-1] + vp[1:]))[:, None], nph, 1),
                               rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------- samplers ----
def _field(rows, nph=8):
    return np.array([[10.0 * j + i for i in range(nph)] for j in range(rows)])


@pytest.mark.parametrize("kind,rows,phi,theta,expect", [
    # nTheta=4, nPhi=8, h=pi/4; hand-derived from KaminoCore.cu:11-184.
    (O.VPHI, 4, 2.75, 0.4, 5.05),      # just north of row 0: flipped pole branch, antipodal mix
    (O.VPHI, 4, 2.75, 0.6, 4.25),      # just south of row 0: ordinary bilinear
    (O.CENTERED, 4, 2.25, 0.3, 5.85),  # flipped pole branch with alpha/2
    (O.VTHETA, 3, 1.5, 3.6, 22.7),     # southern pole row nTheta-2 with alpha/2
    (O.VPHI, 4, 0.25, 3.9, 32.35),     # southern pole row nTheta-1 with full alpha
    (O.VPHI, 4, 0.25, 4.3, 33.95),     # past pi before the offset: no fold
    (O.VPHI, 4, 0.25, 4.7, 31.55),     # folded at pi: phi + pi, flipped
])
def test_sampler_hand_cases(kind, rows, phi, theta, expect):
    h = np.pi / 4
    f = _field(rows)
    v, esc = O.sample(kind, f, 4, 8, phi * h, theta * h)
    assert esc == 0
    assert v == pytest.approx(expect, abs=1e-12)


def test_sampler_escape_is_clamped_and_counted():
    h = np.pi / 4
    f = _field(3)
    v, esc = O.sample(O.VTHETA, f, 4, 8, 0.0, 3.9 * h + h)  # shifted index nTheta-1: past the u_theta grid
    assert esc >= 1 and np.isfinite(v)


def test_validate_coord_folds():
    phi, theta, r = O.validate_coord(0.3, -0.1)
    assert r == -1.0 and theta == pytest.approx(0.1) and phi == pytest.approx(0.3 + np.pi)
    phi, theta, r = O.validate_coord(6.0, np.pi + 0.2)
    assert r == -1.0 and theta == pytest.approx(np.pi - 0.2) and phi == pytest.approx(6.0 + np.pi - 2 * np.pi)


# ------------------------------------------------------------ advection ----
def _rand_state(nt, seed):
    up, ut, _ = O.init_fields(nt, 2 * nt, 6, seed, 0)
    rng = np.random.default_rng(seed)
    return up, ut, rng.uniform(0, 1, (nt, 2 * nt))


def test_zero_velocity_is_identity():
    nt = 16
    _, _, rho = _rand_state(nt, 1)
    up = np.zeros((nt, 2 * nt))
    ut = np.zeros((nt - 1, 2 * nt))
    rng = np.random.default_rng(5)
    up_f = rng.standard_normal(up.shape)
    # zero u_theta and u_phi fields advect themselves and rho unchanged
    op, ot, orr, *_ = O.advect(nt, 2 * nt, 0.01, up, ut, rho)
    # radian node coordinates normalise to integers only up to rounding
    np.testing.assert_allclose(orr, rho, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(op, up)
    np.testing.assert_array_equal(ot, ut)
    assert up_f.shape == up.shape


def test_constant_density_stays_constant():
    nt = 32
    up, ut, _ = _rand_state(nt, 2)
    rho = np.full((nt, 2 * nt), 0.7)
    _, _, orr, *_ = O.advect(nt, 2 * nt, 0.005, up, ut, rho)
    np.testing.assert_allclose(orr, 0.7, rtol=1e-14)


def test_phi_shift_equivariance():
    nt, k = 32, 5
    up, ut, rho = _rand_state(nt, 3)
    a = O.advect(nt, 2 * nt, 0.005, up, ut, rho)
    b = O.advect(nt, 2 * nt, 0.005, np.roll(up, k, 1), np.roll(ut, k, 1), np.roll(rho, k, 1))
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_allclose(np.roll(x, k, 1), y, atol=1e-12)
    s1 = O.step(nt, 2 * nt, 0.005, up, ut, rho)
    s2 = O.step(nt, 2 * nt, 0.005, np.roll(up, k, 1), np.roll(ut, k, 1), np.roll(rho, k, 1))
    for x, y in zip(s1[:3], s2[:3]):
        np.testing.assert_allclose(np.roll(x, k, 1), y, atol=1e-11)


def test_uniform_rotation_by_one_cell_shifts_density():
    nt, dt = 32, 0.01
    h = np.pi / nt
    th = (np.arange(nt) + 0.5) * h
    up = np.repeat((np.sin(th) * h / dt)[:, None], 2 * nt, 1)  # one cell of phi per step
    ut = np.zeros((nt - 1, 2 * nt))
    rho = np.random.default_rng(8).uniform(0, 1, (nt, 2 * nt))
    _, ot, orr, *_ = O.advect(nt, 2 * nt, dt, up, ut, rho)
    np.testing.assert_allclose(orr, np.roll(rho, 1, 1), atol=1e-12)
    np.testing.assert_array_equal(ot, ut)


def _linear_density(nt, alpha, beta, gamma):
    """rho = alpha + beta theta + gamma phi at the rho nodes (ih, (j + 1/2) h):
    bilinear interpolation reproduces it exactly away from the phi seam."""
    h = np.pi / nt
    th = (np.arange(nt) + 0.5) * h
    ph = np.arange(2 * nt) * h
    return alpha + beta * th[:, None] + gamma * ph[None, :], th, ph


def test_uniform_flow_departure_point():
    """Uniform u_theta = c and u_phi = U with R != 1: every velocity sample is
    (U, c), so the RK2 departure point of the rho node g is
    (phi_g - U dt / (R sin theta_g), theta_g - c dt / R) (advectionCentered,
    KaminoCore.cu:292-315: the phi metric stays at theta_g), and a density linear in theta
    and phi is advected to its value there.  Pins the sign of both
    displacements and the dt/R and dt/(R sin theta_g) metrics."""
    nt, dt, R, c, U = 32, 0.01, 2.0, 0.8, 1.5
    alpha, beta, gamma = 0.3, 0.5, 0.02
    rho, th, ph = _linear_density(nt, alpha, beta, gamma)
    up = np.full((nt, 2 * nt), U)
    ut = np.full((nt - 1, 2 * nt), c)
    op, ot, orr, *_ = O.advect(nt, 2 * nt, dt, up, ut, rho, radius=R)
    thp = th - c * dt / R
    php = ph[None, :] - U * dt / (R * np.sin(th))[:, None]
    expect = alpha + beta * thp[:, None] + gamma * php
    rows, cols = slice(2, nt - 2), slice(3, 2 * nt - 3)  # interior stencils, clear of the seam
    np.testing.assert_allclose(orr[rows, cols], expect[rows, cols], rtol=0, atol=1e-12)
    # constant velocity fields advect to themselves
    np.testing.assert_allclose(op[rows], U, rtol=1e-13)
    np.testing.assert_allclose(ot[rows], c, rtol=1e-13)


def test_linear_meridional_flow_rk2_rule():
    """u_theta = k (theta - theta0), u_phi = 0: at a node theta_g the reference
    takes u0 = k (theta_g - theta0), the midpoint m = theta_g - M u0 / 2
    (M = dt / R), u(m) = u0 (1 - k M / 2) and the average of the two,
    ubar = u0 (1 - k M / 4) (KaminoCore.cu:209-224, 254-268, 299-315), so the departure point
    is theta_g - M k (theta_g - theta0)(1 - k M / 4); plain midpoint RK2
    (ubar = u(m)) would give 1 - k M / 2.  Interpolation of the linear fields
    is exact on interior rows, so this pins the averaging rule for rho (at
    rho nodes) and for u_theta itself (at u_theta nodes)."""
    nt, dt, R, k = 32, 0.02, 1.5, 3.0
    th0 = np.pi / 2
    h = np.pi / nt
    M = dt / R
    alpha, beta = 0.2, 0.7
    rho, th, _ = _linear_density(nt, alpha, beta, 0.0)
    tht = (np.arange(nt - 1) + 1.0) * h  # u_theta node latitudes
    ut = np.repeat((k * (tht - th0))[:, None], 2 * nt, 1)
    up = np.zeros((nt, 2 * nt))
    op, ot, orr, *_ = O.advect(nt, 2 * nt, dt, up, ut, rho, radius=R)

    def depart(t):
        return t - M * k * (t - th0) * (1.0 - k * M / 4.0)

    rows = slice(3, nt - 3)
    np.testing.assert_allclose(orr[rows], (alpha + beta * depart(th))[rows, None] * np.ones(2 * nt),
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(ot[rows], (k * (depart(tht) - th0))[rows, None] * np.ones(2 * nt),
                               rtol=0, atol=1e-12)
    np.testing.assert_array_equal(op, 0.0)
    # the plain-midpoint value differs by far more than the tolerance
    plain = th - M * k * (th - th0) * (1.0 - k * M / 2.0)
    assert np.abs(beta * (plain - depart(th))[rows]).max() > 1e-5


def test_ambiguity_mask_is_sparse_for_smooth_flow():
    nt = 64
    up, ut, rho = O.init_fields(nt, 2 * nt, 8, 1, 0)
    *_, ap, at, ar, st = O.advect(nt, 2 * nt, 0.005, up, ut, rho, tol=1e-5)
    assert st.pole_escapes == 0
    assert ap.mean() < 1e-3 and at.mean() < 1e-3 and ar.mean() < 1e-3


def test_full_step_stays_finite_and_bounded():
    nt = 64
    up, ut, rho = O.init_fields(nt, 2 * nt, 8, 1, 0)
    for _ in range(5):
        up, ut, rho, st = O.step(nt, 2 * nt, 0.005, up, ut, rho)
        assert st.pole_escapes == 0
    assert np.isfinite(up).all() and np.abs(up).max() < 2.0
    assert rho.min() >= 0.5 - 1e-12 and rho.max() <= 0.6 + 1e-12  # interpolation is convex
