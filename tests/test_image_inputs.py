"""Image inputs (initDensityfromPic, KaminoSolver.cu:243-277; particle colours,
KaminoParticles.cu:62-76): PPM / PGM reading, the horizontal flip and the
OpenCV-style bilinear resize, checked against closed forms.  Host code only."""
import numpy as np
import pytest

import kaminogpu_amd as K


def write_ppm(path, rgb, ascii_=False):
    h, w, _ = rgb.shape
    if ascii_:
        body = " ".join(str(int(v)) for v in rgb.reshape(-1))
        path.write_bytes(f"P3\n# a comment\n{w} {h}\n255\n{body}\n".encode())
    else:
        path.write_bytes(f"P6\n{w} {h}\n255\n".encode() + rgb.astype(np.uint8).tobytes())


def write_pgm(path, grey):
    h, w = grey.shape
    path.write_bytes(f"P5\n{w} {h}\n255\n".encode() + grey.astype(np.uint8).tobytes())


def test_constant_image_gives_constant_density_and_colours(tmp_path):
    nt = 16
    rgb = np.zeros((5, 7, 3), np.uint8)
    rgb[..., 0], rgb[..., 1], rgb[..., 2] = 200, 100, 50  # R, G, B
    write_ppm(tmp_path / "c.ppm", rgb)
    rho = K.image_density(tmp_path / "c.ppm", nt)
    want = np.float32((np.float32(50 / 255) + np.float32(100 / 255) + np.float32(200 / 255)) / 3.0)
    assert rho.shape == (nt, 2 * nt) and np.all(rho == want)
    coords = K.particles_seed(nt, 1.0, 3)
    col = K.image_colors(tmp_path / "c.ppm", nt, coords)
    np.testing.assert_array_equal(col, np.tile(np.float32([100 / 255, 200 / 255, 50 / 255]), (len(coords), 1)))


def test_same_size_image_is_flipped_not_resampled(tmp_path):
    nt = 8
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, (nt, 2 * nt, 3), dtype=np.uint8)
    write_ppm(tmp_path / "a.ppm", rgb)
    rho = K.image_density(tmp_path / "a.ppm", nt)
    f = rgb[:, ::-1, :].astype(np.float64)  # flip around the vertical axis
    B, G, R = (np.float32(f[..., 2] / 255), np.float32(f[..., 1] / 255), np.float32(f[..., 0] / 255))
    np.testing.assert_array_equal(rho, ((B + G + R).astype(np.float32) / np.float64(3.0)).astype(np.float32))


def test_exact_twofold_downscale_averages_blocks(tmp_path):
    """Halving both axes puts every destination pixel at weights (1/2, 1/2) of a
    2x2 block; a 0/255 checkerboard averages to 127.5, which the 22-bit
    fixed-point rounding turns into 128."""
    nt = 8
    yy, xx = np.mgrid[0:2 * nt, 0:4 * nt]
    grey = np.where((yy + xx) % 2 == 0, 0, 255).astype(np.uint8)
    write_pgm(tmp_path / "k.pgm", grey)
    rho = K.image_density(tmp_path / "k.pgm", nt)
    v = np.float32(128 / 255)
    assert np.all(rho == np.float32((v + v + v) / 3.0))


def test_upscaled_ramp_is_monotone_with_clamped_edges(tmp_path):
    """A two-pixel image (0 | 255) stretched to 2 nTheta columns: after the
    flip the values fall monotonically from 255 to 0, with the outer quarter
    of each side clamped to its edge pixel."""
    nt = 16
    grey = np.array([[0, 255]], np.uint8)
    write_pgm(tmp_path / "r.pgm", grey)
    rho = K.image_density(tmp_path / "r.pgm", nt)
    row = rho[0]
    assert np.all(rho == row)  # one source row: every destination row equal
    assert np.all(np.diff(row) <= 0)
    q = (2 * nt) // 4
    assert np.all(row[:q] == 1.0) and np.all(row[-q:] == 0.0)
    # the midpoint of the ramp
    assert abs(row[nt - 1] + row[nt] - 1.0) < 2 / 255


def test_ascii_and_binary_and_grey_agree(tmp_path):
    nt = 6
    rng = np.random.default_rng(2)
    grey = rng.integers(0, 256, (5, 9), dtype=np.uint8)
    rgb = np.repeat(grey[..., None], 3, axis=2)
    write_pgm(tmp_path / "g.pgm", grey)
    write_ppm(tmp_path / "b.ppm", rgb)
    write_ppm(tmp_path / "a.ppm", rgb, ascii_=True)
    a, b, g = (K.image_density(tmp_path / n, nt) for n in ("a.ppm", "b.ppm", "g.pgm"))
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, g)


def test_missing_or_bad_file_is_an_error(tmp_path):
    with pytest.raises(K.KaminoError):
        K.image_density(tmp_path / "none.ppm", 8)
    (tmp_path / "bad.ppm").write_bytes(b"P7\n1 1\n255\n\x00")
    with pytest.raises(K.KaminoError):
        K.image_density(tmp_path / "bad.ppm", 8)
