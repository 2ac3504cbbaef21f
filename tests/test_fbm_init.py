"""The reference's default initial velocity (FBM value-noise curl,
KaminoInitializer.cu:3-133): the oracle's restatement against closed forms,
and the library's host implementation against the oracle (bit for bit, both
being FP32 / FP64 operation sequences of the reference)."""
import math

import numpy as np
import pytest

import kaminogpu_amd as K
from oracle import oracle as O


def f32(x):
    return float(np.float32(x))


def hash_ref(a, b):
    """rand(vec2(a, b)) (KaminoInitializer.cu:124-133) evaluated with Python
    doubles and libm sin: frac(float(sin(43758.5453 float(12.9898 a + 4.1414 b))))."""
    d = f32(float(a) * 12.9898 + float(b) * 4.1414)
    v = f32(math.sin(d * 43758.5453))
    return f32(v - math.floor(v))


def test_hash_matches_independent_evaluation():
    for a in range(-3, 12):
        for b in range(-2, 9):
            assert O.fbm_hash(a, b) == hash_ref(a, b), (a, b)


def test_hash_range_and_spread():
    """frac(sin(.)) of well-spread arguments follows the arcsine law folded into
    [0, 1): mean 1/2 and standard deviation sqrt(3/4 - 2/pi) = 0.3367."""
    v = np.array([O.fbm_hash(a, b) for a in range(100) for b in range(100)])
    assert v.min() >= 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.01
    assert v.std() == pytest.approx(math.sqrt(0.75 - 2 / math.pi), abs=0.005)


def test_fbm_at_origin_is_zero():
    assert O.fbm_hash(0, 0) == 0.0 and O.fbm(0.0, 0.0) == 0.0


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_fbm_on_lattice_lines_is_the_octave_sum(k):
    """At x = 0, y = k/2 every octave i samples the lattice point (0, 2^i k)
    exactly, so FBM = (1 - 0.5) / 2 * sum_i 0.5^i rand(0, 2^i k)."""
    total = sum(0.5 ** i * O.fbm_hash(0.0, 2 ** i * k) for i in range(4))
    assert O.fbm(0.0, 0.5 * k) == pytest.approx(0.25 * total, abs=2e-7)


@pytest.mark.parametrize("y", [0.5, 1.0, 2.5])
def test_fbm_is_piecewise_linear_along_lattice_rows(y):
    """On y = k/2 the y-fractions vanish in every octave, so FBM is linear in x
    between consecutive points of the finest lattice (spacing 0.15 / 8)."""
    step = 0.15 / 8
    for m in (3, 10, 41):
        xs = [(m + f) * step for f in (0.2, 0.5, 0.8)]
        a, b, c = (O.fbm(x, y) for x in xs)
        assert b == pytest.approx(0.5 * (a + c), abs=3e-6)


def test_velocity_is_the_noise_curl_with_the_reference_corners():
    """u_phi from the theta differences of the noise at the four corners of a
    u_phi cell; u_theta from the phi differences, with the lower-left corner
    taken at the upper-left point (KaminoInitializer.cu:69)."""
    nt, R = 12, 1.5
    nph = 2 * nt
    up, ut = O.init_fbm(nt, nph, R)
    h = f32(2 * math.pi / nph)
    h2 = f32(h / 2)
    gain = f32(4096.0 / nph)
    F = O.fbm

    def at(k):  # k h in FP32, as the reference forms the corner coordinates
        return f32(k * h)

    for j in (0, 5, nt - 1):
        for i in (1, 7, nph - 1):
            xr, xl = f32(at(i) + h2), f32(at(i) - h2)
            d = (F(xr, at(j + 1)) - F(xr, at(j))) + (F(xl, at(j + 1)) - F(xl, at(j)))
            assert up[j, i] == pytest.approx(gain * d / (2 * R * h), rel=1e-5, abs=1e-5)
    # seam column: west corners at 2 pi - h / 2
    j, xw = 4, f32(2 * math.pi - h2)
    d = (F(h2, at(j + 1)) - F(h2, at(j))) + (F(xw, at(j + 1)) - F(xw, at(j)))
    assert up[j, 0] == pytest.approx(gain * d / (2 * R * h), rel=1e-5, abs=1e-5)
    for j in (1, 6, nt - 1):
        for i in (0, 9, nph - 1):
            yu, yd = f32(at(j) + h2), f32(at(j) - h2)
            ul = F(at(i), yu)
            d = -(F(at(i + 1), yu) - ul) - (F(at(i + 1), yd) - ul)
            assert ut[j - 1, i] == pytest.approx(gain * d / (2 * R * h), rel=1e-5, abs=1e-5)


@pytest.mark.parametrize("nt", [8, 48, 64])
def test_library_matches_oracle_bitwise(nt):
    up, ut = K.init_fbm(nt, radius=2.0)
    rp, rt = O.init_fbm(nt, 2 * nt, 2.0)
    np.testing.assert_array_equal(up, rp)
    np.testing.assert_array_equal(ut, rt)
