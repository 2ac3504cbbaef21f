"""Whole runs of the benchmark workloads against FP64 oracle trajectories.

Both sides start from the same FP32 state; the oracle then steps in FP64 all
the way (its state is never rounded to FP32), so the reported error is that of
the FP32 path against an FP64 run of the reference's algorithm, accumulated
over the run.  Per field: relative L2 of u_phi, u_theta and rho, and of the
divergence D(u) (the reference's fillDivergenceKernel operator) of the final
velocity.  The per-step bar (1e-5, test_gpu_parity) is the gate; a whole run
is bounded by TOL_RUN, a regression bound, and its values are printed.

The pole rows dominate a whole run's error.  The reference's pole samplers
are discontinuous (a departure point just past row 0's latitude samples the
antipodal value), so two trajectories that differ by rounding part there at
some step: the FP64 oracle itself, perturbed each step by random noise of
3e-6 of each velocity field's maximum, differs from its unperturbed run by
4.7e-2 of the u_theta maximum in row 0 after 20 steps at 1024x2048.  The
error is therefore also reported, and bounded more tightly, over the
interior rows (INTERIOR rows and more away from either pole).

One 256x512 sphere, dt 0.002, all 1000 steps.  One 1024x2048 sphere, dt
0.5/1024, the first 100 of its 500 steps (the FP64 oracle's 500 steps take
minutes on the host).  The uneven batch: all 64 spheres over all 200 steps, stepped together by the engine from 32-step
hipGraphs in three concurrent groups, as the benchmark runs them.
"""
import numpy as np
import pytest

import kaminogpu_amd as K
from oracle import oracle as O
from parity_util import f64, log, rel_l2, seeded_state

pytestmark = pytest.mark.gpu

TOL_RUN = 2e-3       # relative L2 of a field after a whole run (regression bound)
TOL_INTERIOR = 1e-4  # the same over the interior rows
INTERIOR = 16        # rows nearest each pole left out of the interior


def oracle_run(nt, dt, state, steps):
    up, ut, rho = [np.ascontiguousarray(a, np.float64) for a in f64(*state)]
    O.run(nt, 2 * nt, dt, up, ut, rho, steps)
    return up, ut, rho


def run_errors(nt, got, ref):
    """[u_phi, u_theta, rho, div] relative L2 over the whole field, and the
    same three fields over the interior rows."""
    errs = [rel_l2(g, r) for g, r in zip(got, ref)]
    d_got = O.divergence(nt, 2 * nt, *f64(got[0], got[1]))
    d_ref = O.divergence(nt, 2 * nt, ref[0], ref[1])
    errs.append(float(np.linalg.norm(d_got - d_ref) / np.linalg.norm(d_ref)))
    k = min(INTERIOR, nt // 4)
    inner = [rel_l2(g[k:g.shape[0] - k], r[k:r.shape[0] - k]) for g, r in zip(got, ref)]
    return errs, inner


@pytest.mark.parametrize("name,nt,dt,L,steps", [("s256", 256, 0.002, 16, 1000),
                                                ("s1024", 1024, 0.5 / 1024, 32, 100)])
def test_trajectory_single_sphere(name, nt, dt, L, steps):
    state = seeded_state(nt, L=L, seed=1, dens=1)
    log(f"[run] {name} {nt}x{2 * nt} {steps} steps: FP64 oracle trajectory")
    ref = oracle_run(nt, dt, state, steps)
    for prec, tag in ((K.PREC_FP32, "p32"), (K.PREC_FP64_PRESSURE, "p64")):
        with K.Sim(nt, dt, precision=prec) as s:
            s.set_fields(*state)
            s.step(steps)
            got = s.get_fields()
        errs, inner = run_errors(nt, got, ref)
        log(f"[run] {name} {tag} after {steps} steps rel L2: u_phi {errs[0]:.2e} u_theta {errs[1]:.2e} "
            f"rho {errs[2]:.2e} div {errs[3]:.2e}; interior u_phi {inner[0]:.2e} u_theta {inner[1]:.2e} "
            f"rho {inner[2]:.2e}")
        assert max(errs[:3]) < TOL_RUN
        assert max(inner) < TOL_INTERIOR


def test_trajectory_uneven_batch():
    rng = np.random.default_rng(3)
    sizes = [int(n) for n in rng.choice([32, 48, 64, 96, 128, 192, 256], 64)]
    steps = 200
    states = [seeded_state(nt, L=min(16, nt // 4), seed=100 + k, dens=1) for k, nt in enumerate(sizes)]
    with K.Engine(use_graph=True, graph_steps=32, groups=3) as eng:
        ids = [eng.admit(nt, 0.5 / nt, *st, steps=steps) for nt, st in zip(sizes, states)]
        assert all(i is not None for i in ids)
        eng.step(steps)
        got = [eng.fetch(i, nt) for i, nt in zip(ids, sizes)]
    worst = {}
    for nt, st, g in zip(sizes, states, got):
        errs, inner = run_errors(nt, g, oracle_run(nt, 0.5 / nt, st, steps))
        w = worst.setdefault(nt, [0.0] * 7)
        worst[nt] = [max(a, b) for a, b in zip(w, errs + inner)]
    for nt in sorted(worst):
        e = worst[nt]
        log(f"[run] uneven n_theta={nt} worst after {steps} steps rel L2: u_phi {e[0]:.2e} u_theta {e[1]:.2e} "
            f"rho {e[2]:.2e} div {e[3]:.2e}; interior u_phi {e[4]:.2e} u_theta {e[5]:.2e} rho {e[6]:.2e}")
    assert max(max(e[:3]) for e in worst.values()) < TOL_RUN
    assert max(max(e[4:]) for e in worst.values()) < TOL_INTERIOR


def test_fp64_pressure_pole_rows():
    """One step at 1024x2048: the pressure gradient divides differences of p
    by sin(theta), 1/651 at row 0, so p rounded to FP32 leaves row 0 of u_phi
    ~2e-5 of the field's maximum from the oracle; with the FP64 pressure every
    row stays at FP32 rounding."""
    nt, dt = 1024, 0.5 / 1024
    state = seeded_state(nt, L=32, seed=1, dens=1)
    ref = O.step(nt, 2 * nt, dt, *f64(*state))
    worst = {}
    for prec, tag in ((K.PREC_FP32, "p32"), (K.PREC_FP64_PRESSURE, "p64")):
        with K.Sim(nt, dt, precision=prec) as s:
            s.set_fields(*state)
            s.step(1)
            got = s.get_fields()
        fm = np.abs(ref[0]).max()
        row = np.abs(got[0] - ref[0]).max(axis=1) / fm
        worst[tag] = (row[0], row[nt - 1], row[4:nt - 4].max())
        log(f"[pole] {tag} one step, u_phi max error / field max: row 0 {row[0]:.2e}, row {nt - 1} "
            f"{row[nt - 1]:.2e}, rows 4..{nt - 5} {row[4:nt - 4].max():.2e}")
    assert max(worst["p64"]) < 3e-7
    assert worst["p32"][0] > 10 * worst["p64"][0]
