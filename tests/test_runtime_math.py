"""CPU checks of integer identities the kernels rely on."""
import numpy as np


def test_magic_division_exact_for_every_supported_width():
    # k_advect splits a flat item index into (row, column) as
    # row = umulhi(idx, ceil(2^32 / nPhi)); items stay below 2^16.
    idx = np.arange(1 << 16, dtype=np.uint64)
    for n_phi in range(8, 2 * 1024 + 1, 2):
        m = ((1 << 32) + n_phi - 1) // n_phi
        q = (idx * np.uint64(m)) >> np.uint64(32)
        assert np.array_equal(q, idx // np.uint64(n_phi)), n_phi


def test_advection_items_fit_the_magic_range():
    # the largest flat loop: (R + 2) rows of nPhi with R <= 8, nPhi <= 2048
    assert (8 + 2) * 2048 < (1 << 16)
