"""CPU checks of integer identities the kernels rely on."""
import numpy as np


def test_magic_division_exact_for_every_divisor_used():
    # Flat item loops split an index as q = umulhi(i, ceil(2^32 / d)): tile
    # widths W, W+1, W+2 and the staged width (k_advect), K = rows per lane
    # (k_tridiag staging), rows per FFT tile, the Stockham stride Ls of a
    # row-FFT pass when it is not a power of two; every divisor 2..2048
    # (K = 1 and R = 1 are taken apart: their multiplier 2^32 does not fit
    # 32 bits), every dividend below 2^16.
    idx = np.arange(1 << 16, dtype=np.uint64)
    for d in range(2, 2049):
        m = ((1 << 32) + d - 1) // d
        q = (idx * np.uint64(m)) >> np.uint64(32)
        assert np.array_equal(q, idx // np.uint64(d)), d


def test_flat_loops_fit_the_magic_range():
    # k_advect: staged block (R + 6)(W + 8) and u_phi items (R + 1)(W + 2) with
    # R <= 64, W <= 128; k_tridiag staging: (256 / L) modes x nTheta <= 4096
    assert (64 + 6) * (128 + 8) < (1 << 16)
    assert (64 + 1) * (128 + 2) < (1 << 16)
    assert max((256 // L) * nt for L, nt in [(8, 64), (16, 128), (32, 256), (64, 1024)]) < (1 << 16)


def test_32_bit_form_of_the_magic_constant():
    # k_ifft_grad forms ceil(2^32 / d) on the device as 0xffffffff / d + 1
    # (d = rows + 1 >= 2): equal for every divisor, powers of two included
    for d in range(2, 2049):
        assert 0xFFFFFFFF // d + 1 == ((1 << 32) + d - 1) // d, d
