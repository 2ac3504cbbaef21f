"""GPU parity of the phi transform as dense DFT products on MFMA (FP16 hi/lo
operands, FP32 accumulate), the autotuner's alternative to the row FFT:
against the FP64 oracle at the per-field bar, through the same C ABI."""
import numpy as np
import pytest

import kaminogpu_amd as K
from oracle import oracle as O
from parity_util import TOL_FIELD, f32, f64, log, oracle_advect_geo, rel_l2, seeded_state

pytestmark = pytest.mark.gpu


class dft_for:
    """Selects the MFMA DFT for simulations of these sizes while in scope."""

    def __init__(self, *sizes):
        self.sizes = sizes

    def __enter__(self):
        for nt in self.sizes:
            K.tune_set_transform(nt, K.TRANSFORM_DFT)

    def __exit__(self, *a):
        for nt in self.sizes:
            K.tune_set_transform(nt, K.TRANSFORM_FFT)


def projected(nt, dt, state):
    with K.Sim(nt, dt) as s:
        s.set_fields(*state)
        s.step_phases(K.PHASE_PROJECT)
        return s.get_fields()


@pytest.mark.parametrize("nt", [8, 32, 45, 64, 75, 192, 256, 512, 1024])
def test_dft_projection_parity(nt):
    dt = 0.5 / nt
    state = seeded_state(nt, L=max(1, min(16, nt // 4)), seed=5 * nt, dens=1)
    (up, ut, _), _, _ = oracle_advect_geo(nt, dt, *state)
    up, ut = f32(up, ut)
    with dft_for(nt):
        assert K.tune_get_transform(nt) == K.TRANSFORM_DFT
        g = projected(nt, dt, (up, ut, state[2]))
    fft = projected(nt, dt, (up, ut, state[2]))
    r_up, r_ut, _, _ = O.project(nt, 2 * nt, *f64(up, ut))
    e = (rel_l2(g[0], r_up), rel_l2(g[1], r_ut))
    e_fft = (rel_l2(fft[0], r_up), rel_l2(fft[1], r_ut))
    log(f"[dft project] nt={nt} rel L2 u_phi={e[0]:.2e} u_theta={e[1]:.2e} "
        f"(FFT path {e_fft[0]:.2e} {e_fft[1]:.2e})")
    np.testing.assert_array_equal(g[2], state[2])
    assert max(e) < TOL_FIELD


@pytest.mark.parametrize("nt", [48, 256])
def test_dft_spectrum_matches_oracle(nt):
    nph, M = 2 * nt, nt
    dt = 0.5 / nt
    state = seeded_state(nt, L=min(16, nt // 4), seed=7, dens=1)
    (up, ut, _), _, _ = oracle_advect_geo(nt, dt, *state)
    up, ut = f32(up, ut)
    with dft_for(nt):
        with K.Sim(nt, dt) as s:
            s.set_fields(up, ut, state[2])
            s.step_phases(K.PHASE_DIVFFT)
            dhat = s.spectrum()
    _, _, spec, _ = O.project(nt, nph, *f64(up, ut))
    ref = spec[nph // 2:nph // 2 + M, :]
    nyq = spec[0, :].real
    got_nyq = dhat[0].imag.copy()
    got = dhat.copy()
    got[0] = got[0].real
    e = rel_l2(np.concatenate([got.real.ravel(), got.imag.ravel(), got_nyq]),
               np.concatenate([ref.real.ravel(), ref.imag.ravel(), nyq]))
    log(f"[dft spectrum] nt={nt} rel L2 d-hat={e:.2e}")
    assert e < TOL_FIELD


def test_dft_full_step_config2():
    """Config 2's sphere, 1024x2048, one full step with the MFMA transform
    against the oracle (ambiguous cells substituted as in the FFT test)."""
    nt, dt = 1024, 0.5 / 1024
    state = seeded_state(nt, L=32, seed=1, dens=1, pre_steps=2, dt=dt)
    with dft_for(nt):
        with K.Sim(nt, dt) as s:
            s.set_fields(*state)
            s.step_phases(K.PHASE_ADVECT)
            g_adv = s.get_fields()
            s.step_phases(K.PHASE_PROJECT)
            got = s.get_fields()
    (a_up, a_ut, a_rho), masks, st = oracle_advect_geo(nt, dt, *state)
    a_up = np.where(masks[0] != 0, g_adv[0], a_up)
    a_ut = np.where(masks[1] != 0, g_adv[1], a_ut)
    r_up, r_ut, _, _ = O.project(nt, 2 * nt, *f64(a_up, a_ut))
    e = (rel_l2(got[0], r_up), rel_l2(got[1], r_ut), rel_l2(got[2], a_rho, masks[2]))
    log(f"[dft step] 1024x2048 rel L2 u_phi={e[0]:.2e} u_theta={e[1]:.2e} rho={e[2]:.2e}")
    assert max(e) < TOL_FIELD


def test_dft_pressure_readback():
    nt = 64
    rng = np.random.default_rng(3)
    up = rng.uniform(-1, 1, (nt, 2 * nt)).astype(np.float32)
    ut = rng.uniform(-1, 1, (nt - 1, 2 * nt)).astype(np.float32)
    with dft_for(nt):
        with K.Sim(nt, 0.01) as s:
            s.set_fields(up, ut, np.zeros((nt, 2 * nt), np.float32))
            s.step_phases(K.PHASE_PROJECT)
            p = s.pressure()
    _, _, _, r_p = O.project(nt, 2 * nt, *f64(up, ut))
    e = rel_l2(p, r_p)
    log(f"[dft pressure] nt={nt} rel L2 {e:.2e}")
    assert e < TOL_FIELD


def test_dft_and_fft_simulations_in_one_batch_equal_alone():
    """A batch mixing MFMA-DFT sizes, FFT sizes and an FP64-pressure
    simulation of a DFT size (which keeps the FFT) equals solo runs bitwise."""
    sizes = [32, 48, 64, 75, 64]
    precs = [K.PREC_FP32, K.PREC_FP32, K.PREC_FP32, K.PREC_FP32, K.PREC_FP64_PRESSURE]
    reqs = [(nt, seeded_state(nt, L=min(16, nt // 4), seed=60 + k, dens=1)) for k, nt in enumerate(sizes)]
    with dft_for(32, 64, 75):
        with K.Engine() as eng:
            ids = [eng.admit(nt, 0.5 / nt, *st, steps=4, precision=p) for (nt, st), p in zip(reqs, precs)]
            eng.step(4)
            got = [eng.fetch(i, nt) for (nt, _), i in zip(reqs, ids)]
        alone = []
        for (nt, st), p in zip(reqs, precs):
            with K.Sim(nt, 0.5 / nt, precision=p) as s:
                s.set_fields(*st)
                s.step(4)
                alone.append(s.get_fields())
    for g, a in zip(got, alone):
        for x, y in zip(g, a):
            np.testing.assert_array_equal(x, y)


def test_autotune_records_the_transform(tmp_path):
    cache = str(tmp_path / "tune.txt")
    ms = K.autotune(48, 16, steps=3, cache_path=cache)
    line = open(cache).read().split()
    log(f"[autotune] 48 -> {line}, kernel ms {np.round(ms, 4)}")
    assert len(line) == 9 and int(line[8]) in (K.TRANSFORM_FFT, K.TRANSFORM_DFT, K.TRANSFORM_FFT8)
    assert K.tune_get_transform(48) == int(line[8])
    K.tune_set_transform(48, K.TRANSFORM_FFT)
    assert 48 in K.tune_load(cache) and K.tune_get_transform(48) == int(line[8])
    K.tune_set_transform(48, K.TRANSFORM_FFT)


def test_eight_wave_forward_transform_equals_the_row_fft():
    """The forward transform of a batch whose FFT simulations all ask for
    TRANSFORM_FFT8 (half twiddle table, eight waves per SIMD) gives the same
    fields, bitwise, as the default kernel: the table's second half is the
    exact negation of its first.  A batch mixing the two takes the default."""
    sizes = [45, 64, 256, 512, 1024]
    reqs = [(nt, seeded_state(nt, L=min(16, nt // 4), seed=80 + k, dens=1)) for k, nt in enumerate(sizes)]
    out = []
    for t8 in ([], sizes, sizes[:2]):
        for nt in t8:
            K.tune_set_transform(nt, K.TRANSFORM_FFT8)
        try:
            with K.Engine() as eng:
                ids = [eng.admit(nt, 0.5 / nt, *st, steps=3) for nt, st in reqs]
                eng.step(3)
                out.append([eng.fetch(i, nt) for (nt, _), i in zip(reqs, ids)])
        finally:
            for nt in sizes:
                K.tune_set_transform(nt, K.TRANSFORM_FFT)
    for run in out[1:]:
        for g, a in zip(run, out[0]):
            for x, y in zip(g, a):
                np.testing.assert_array_equal(x, y)
    log(f"[fft8] sizes {sizes}: eight-wave forward transform bitwise equal to the default")
