"""GPU tests of batching and serving: every simulation's result is the same
alone (kam_sim, plain launches) and inside any batch or serving mix (engine,
graph-replayed batched launches).  A simulation's tiles do not depend on what
else is in the batch, so the equality is bitwise."""
import numpy as np
import pytest

import kaminogpu_amd as K
from oracle import oracle as O
from parity_util import TOL_FIELD, f64, log, oracle_advect_geo, rel_l2, seeded_state

pytestmark = pytest.mark.gpu

SIZES = [32, 48, 64, 96, 128, 160, 192, 256]


def alone(nt, dt, state, steps, geometric=True, precision=K.PREC_FP32):
    with K.Sim(nt, dt, geometric=geometric, precision=precision) as s:
        s.set_fields(*state)
        s.step(steps)
        return s.get_fields()


def make_request(nt, seed):
    return dict(nt=nt, dt=0.5 / nt, state=seeded_state(nt, L=min(16, nt // 4), seed=seed, dens=1))


def test_uneven_batch_equals_alone_bitwise():
    reqs = [make_request(nt, 100 + k) for k, nt in enumerate(SIZES)]
    with K.Engine() as eng:
        ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=6) for r in reqs]
        assert all(i is not None for i in ids)
        eng.step(6)
        eng.sync()
        for r, i in zip(reqs, ids):
            assert eng.poll(i).state == K.engine.REQ_DONE
            got = eng.fetch(i, r["nt"])
            ref = alone(r["nt"], r["dt"], r["state"], 6)
            for g, x in zip(got, ref):
                np.testing.assert_array_equal(g, x)
    log(f"[engine] uneven batch of {len(reqs)} sims bitwise equal to alone")


def test_mixed_precision_batch_equals_alone_bitwise():
    """FP32- and FP64-pressure simulations of several sizes in one batch (the
    inverse transform goes out as one launch per precision): each equals its
    solo run bitwise."""
    sizes = [48, 64, 75, 128, 256, 375]
    reqs = [make_request(nt, 500 + k) for k, nt in enumerate(sizes)]
    precs = [K.PREC_FP64_PRESSURE if k % 2 else K.PREC_FP32 for k in range(len(sizes))]
    with K.Engine() as eng:
        ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=5, precision=p) for r, p in zip(reqs, precs)]
        assert all(i is not None for i in ids)
        eng.step(5)
        got = [eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)]
    for r, p, g in zip(reqs, precs, got):
        for a, b in zip(g, alone(r["nt"], r["dt"], r["state"], 5, precision=p)):
            np.testing.assert_array_equal(a, b)
    log(f"[engine] batch of FP32- and FP64-pressure sims {sizes} bitwise equal to alone")


def test_mixed_tile_widths_equal_alone_bitwise():
    # 256 tuned to 128-column tiles, 128 to 64 and 64 to 32: alone, each stages
    # at its own stride; in the batch 256 is restaged as 64-column tiles and
    # every tile is staged at the 64-column stride
    sizes = [256, 128, 64]
    reqs = [make_request(nt, 300 + k) for k, nt in enumerate(sizes)]
    try:
        K.tune_set(256, 8, 128, 4)
        K.tune_set(128, 16, 64, 4)
        K.tune_set(64, 16, 32, 4)
        with K.Engine() as eng:
            ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=4) for r in reqs]
            eng.step(4)
            got = [eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)]
        for r, g in zip(reqs, got):
            for a, b in zip(g, alone(r["nt"], r["dt"], r["state"], 4)):
                np.testing.assert_array_equal(a, b)
    finally:
        K.tune_set(256, 0, 0, 0)
        K.tune_set(128, 0, 0, 0)
        K.tune_set(64, 0, 0, 0)
    log("[engine] batch mixing 128-, 64- and 32-column tiles bitwise equal to alone")


def test_graph_and_direct_launches_agree():
    reqs = [make_request(nt, 7 + k) for k, nt in enumerate([64, 96, 64])]
    out = []
    for use_graph in (True, False):
        with K.Engine(use_graph=use_graph) as eng:
            ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=5) for r in reqs]
            eng.step(5)
            out.append([eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)])
    for a, b in zip(*out):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_multi_step_graphs_equal_direct_launches():
    """Runs of 8 steps captured as one graph: requests finishing at step counts
    that are not multiples of 8, a request admitted mid-run and odd call
    lengths give the same fields, bitwise, as direct launches, and the step
    accounting is unchanged."""
    reqs = [make_request(nt, 500 + k) for k, nt in enumerate([64, 45, 96, 128])]
    steps = [21, 8, 13, 30]
    late = make_request(75, 520)
    out, counts = [], []
    for graph in (True, False):
        with K.Engine(use_graph=graph, graph_steps=8) as eng:
            ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=s) for r, s in zip(reqs, steps)]
            eng.step(11)
            lid = eng.admit(late["nt"], late["dt"], *late["state"], steps=19)
            eng.step(9)
            eng.step(17)
            eng.step(3)
            st = eng.stats()
            counts.append((st.steps, st.cell_steps, st.completed, st.active))
            got = [eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)]
            got.append(eng.fetch(lid, late["nt"]))
            out.append(got)
    assert counts[0] == counts[1] and counts[0][2] == 5 and counts[0][3] == 0, counts
    for a, b in zip(*out):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    log(f"[engine] 8-step graphs = direct launches bitwise; steps {counts[0][0]}")


@pytest.mark.parametrize("groups,graph", [(2, True), (2, False), (3, True)])
def test_concurrent_groups_equal_one_group(groups, graph):
    """Simulations split into groups stepped on streams of their own give the
    same fields, bitwise, and the same step accounting as one group: sizes of
    every kind, both precisions, retirements at step counts that are not
    multiples of the graph length, and admissions mid-run (which rebuild every
    group)."""
    sizes = [256, 64, 45, 96, 128, 75, 192]
    steps = [21, 8, 13, 30, 17, 9, 26]
    precs = [K.PREC_FP64_PRESSURE if k in (1, 4) else K.PREC_FP32 for k in range(len(sizes))]
    reqs = [make_request(nt, 600 + k) for k, nt in enumerate(sizes)]
    late = [make_request(48, 620), make_request(160, 621)]
    out, counts = [], []
    for g in (1, groups):
        with K.Engine(use_graph=graph, graph_steps=8, groups=g) as eng:
            ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=s, precision=p)
                   for r, s, p in zip(reqs, steps, precs)]
            eng.step(11)
            lids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=19) for r in late]
            eng.step(9)
            eng.step(17)
            st = eng.stats()
            counts.append((st.steps, st.cell_steps, st.completed, st.active))
            got = [eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)]
            got += [eng.fetch(i, r["nt"]) for r, i in zip(late, lids)]
            out.append(got)
    assert counts[0] == counts[1] and counts[0][2] == 9 and counts[0][3] == 0, counts
    for a, b in zip(*out):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    log(f"[engine] {groups} concurrent groups (graph={graph}) bitwise equal to one group")


def test_one_batched_step_matches_oracle():
    """One shared step of a batch spanning the serving mix's sizes (odd n_theta,
    the 64-lane tridiagonal class up to 512): each simulation equals its solo
    run bitwise, and the oracle's step where no cell is branch-ambiguous (the
    solo path is compared with the oracle cell by cell, ambiguous cells
    included, in test_full_step_parity)."""
    sizes = [45, 48, 64, 75, 128, 192, 270, 375, 512]
    reqs = [make_request(nt, 40 + k) for k, nt in enumerate(sizes)]
    with K.Engine() as eng:
        ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=1) for r in reqs]
        eng.step(1)
        for r, i in zip(reqs, ids):
            nt, dt = r["nt"], r["dt"]
            got = eng.fetch(i, nt)
            for g, x in zip(got, alone(nt, dt, r["state"], 1)):
                np.testing.assert_array_equal(g, x)
            ref = O.step(nt, 2 * nt, dt, *f64(*r["state"]))
            _, masks, _ = oracle_advect_geo(nt, dt, *r["state"])
            amb = [int(m.sum()) for m in masks]
            errs = [rel_l2(got[0], ref[0]), rel_l2(got[1], ref[1]), rel_l2(got[2], ref[2], masks[2])]
            log(f"[engine] nt={nt} one step rel L2 {['%.2e' % e for e in errs]} ambiguous={amb}")
            if amb[0] + amb[1] == 0:
                assert max(errs) < TOL_FIELD
            assert errs[2] < TOL_FIELD


@pytest.mark.parametrize("use_graph,groups", [(False, 1), (True, 1), (False, 3)])
def test_serving_mix_continuous_admission(use_graph, groups):
    """Requests arrive while others run; each is admitted as a slot frees up,
    runs its own step count inside shared steps, and equals its solo run
    (with graph replay, every membership change re-captures the step)."""
    rng = np.random.default_rng(5)
    pending = []
    for k in range(10):
        nt = int(rng.choice([32, 48, 64, 96]))
        r = make_request(nt, 300 + k)
        r["steps"] = int(rng.integers(2, 9))
        r["arrive"] = int(rng.integers(0, 12))
        pending.append(r)
    pending.sort(key=lambda r: r["arrive"])
    done = {}
    with K.Engine(max_slots=3, use_graph=use_graph, groups=groups) as eng:
        running = {}
        t = 0
        while pending or running:
            while pending and pending[0]["arrive"] <= t:
                rid = eng.admit(pending[0]["nt"], pending[0]["dt"], *pending[0]["state"],
                                steps=pending[0]["steps"])
                if rid is None:
                    break  # no slot now; retry after the next step
                running[rid] = pending.pop(0)
            eng.step(1)
            t += 1
            for rid in list(running):
                st = eng.poll(rid)
                if st.state == K.engine.REQ_DONE:
                    r = running.pop(rid)
                    done[id(r)] = (r, eng.fetch(rid, r["nt"]))
            assert t < 500
        s = eng.stats()
        assert s.completed == 10 and s.active == 0
    for r, got in done.values():
        ref = alone(r["nt"], r["dt"], r["state"], r["steps"])
        for g, x in zip(got, ref):
            np.testing.assert_array_equal(g, x)
    log(f"[serve] 10 requests through 3 slots in {t} shared steps, all equal to solo runs")


def test_graph_steps_bounds():
    """graph_steps above 64 is refused at creation; 0 means one step per graph."""
    with pytest.raises(K.KaminoError, match="graph_steps"):
        K.Engine(use_graph=True, graph_steps=65)
    r = make_request(32, 9)
    with K.Engine(use_graph=True, graph_steps=0) as eng:
        i = eng.admit(r["nt"], r["dt"], *r["state"], steps=2)
        eng.step(2)
        got = eng.fetch(i, r["nt"])
    for a, b in zip(got, alone(r["nt"], r["dt"], r["state"], 2)):
        np.testing.assert_array_equal(a, b)


def test_groups_steps_around_a_profile():
    """Steps timed per kernel (kam_engine_profile, every group on the engine
    stream) between ordinary steps of three groups: the fields equal those of
    one group doing the same, bitwise."""
    reqs = [make_request(nt, 40 + k) for k, nt in enumerate([128, 64, 96, 45, 75])]
    out = []
    for groups in (1, 3):
        with K.Engine(use_graph=True, graph_steps=4, groups=groups) as eng:
            ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=20) for r in reqs]
            eng.step(5)
            eng.profile(3)
            eng.step(12)
            out.append([eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)])
    for a, b in zip(*out):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_groups_bounds_and_empty_groups():
    """groups above 4 is refused; with more groups than simulations the empty
    groups launch nothing and the results equal solo runs."""
    with pytest.raises(K.KaminoError, match="groups"):
        K.Engine(groups=5)
    reqs = [make_request(nt, 30 + k) for k, nt in enumerate([64, 45])]
    for graph in (True, False):
        with K.Engine(use_graph=graph, graph_steps=4, groups=4) as eng:
            ids = [eng.admit(r["nt"], r["dt"], *r["state"], steps=6) for r in reqs]
            eng.step(6)
            got = [eng.fetch(i, r["nt"]) for r, i in zip(reqs, ids)]
        for r, g in zip(reqs, got):
            for a, b in zip(g, alone(r["nt"], r["dt"], r["state"], 6)):
                np.testing.assert_array_equal(a, b)


def test_hbm_budget_admission_and_cancel():
    nt = 64
    r = make_request(nt, 1)
    per = K.lib().kam_sim_bytes(nt, 2 * nt)
    with K.Engine(hbm_budget_bytes=2 * per + 4096) as eng:
        a = eng.admit(nt, r["dt"], *r["state"], steps=100)
        b = eng.admit(nt, r["dt"], *r["state"], steps=100)
        assert a is not None and b is not None
        assert eng.admit(nt, r["dt"], *r["state"], steps=100) is None  # budget exhausted
        eng.step(2)
        eng.cancel(a)
        c = eng.admit(nt, r["dt"], *r["state"], steps=3)
        assert c is not None
        eng.step(3)
        eng.sync()  # poll reports ACTIVE until the final copy to the host has landed
        assert eng.poll(c).state == K.engine.REQ_DONE
        got = eng.fetch(c, nt)
        ref = alone(nt, r["dt"], r["state"], 3)
        for g, x in zip(got, ref):
            np.testing.assert_array_equal(g, x)
        s = eng.stats()
        assert s.rejected == 1 and s.cancelled == 1 and s.active == 1


def test_engine_stats_latency():
    reqs = [make_request(256, 9 + k) for k in range(4)]
    with K.Engine() as eng:
        for r in reqs:
            eng.admit(256, r["dt"], *r["state"], steps=1000)
        eng.step(5)
        eng.reset_stats()
        eng.step(50)
        s = eng.stats()
        assert s.steps == 50 and s.cell_steps == 50 * 4 * 256 * 512
        assert 0 < s.step_p50_ms <= s.step_p99_ms <= s.step_max_ms
        k = eng.profile(3)
        assert (k > 0).all()
        log(f"[engine] 4x256x512: p50 {s.step_p50_ms:.3f} ms p99 {s.step_p99_ms:.3f} ms "
            f"kernels {np.round(k, 4)}")


def test_poisson_serving_stream_completes_and_matches_solo():
    """A short Poisson stream through serve() over the serving mix's full size
    range (n_theta log-uniform in [32, 512], odd sizes included): every request
    completes and the completion query reports each once.  Then the same
    requests run again through 6 slots with continuous admission, and every
    result is compared with its solo run."""
    from kaminogpu_amd import serve

    reqs = serve.make_requests(24, seed=11, lo=32, hi=512, min_steps=5, max_steps=40)
    sizes = {r.n_theta for r in reqs}
    assert any(n % 2 for n in sizes) and any(n > 256 for n in sizes), sizes
    pool = serve.StatePool(L=4, variants=2)
    with K.Engine(max_slots=6) as eng:
        out = serve.serve(reqs, rate=400.0, engine=eng, pool=pool)
    assert out["requests"] == 24
    assert all(r.done >= r.admit >= r.arrive for r in reqs)
    assert out["request_p99_s"] >= out["request_p50_s"] > 0
    log(f"[serve] 24 requests, n_theta {sorted(sizes)}: {out['cell_steps_per_s']:.3e} cell-steps/s, "
        f"request p50 {out['request_p50_s'] * 1e3:.1f} ms p99 {out['request_p99_s'] * 1e3:.1f} ms")
    pending = list(reqs)
    running, results = {}, {}
    with K.Engine(max_slots=6) as eng:
        steps = 0
        while pending or running:
            while pending:
                r = pending[0]
                rid = eng.admit(r.n_theta, 0.5 / r.n_theta, *pool.get(r.n_theta, r.seed), steps=r.steps)
                if rid is None:
                    break
                running[rid] = pending.pop(0)
            eng.step(1)
            steps += 1
            for rid in eng.completed():
                r = running.pop(rid)
                results[rid] = (r, eng.fetch(rid, r.n_theta))
            assert steps < 2000
    assert len(results) == 24
    for r, got in results.values():
        ref = alone(r.n_theta, 0.5 / r.n_theta, pool.get(r.n_theta, r.seed), r.steps)
        for g, x in zip(got, ref):
            np.testing.assert_array_equal(g, x)
    log(f"[serve] all 24 results equal their solo runs ({steps} shared steps)")


def test_nan_quarantines_only_that_simulation():
    """A non-finite value injected into one simulation's state sets its
    watchdog word; the engine retires it as FAILED within 8 steps and every
    other simulation runs on, equal to its solo run."""
    reqs = [make_request(nt, 60 + k) for k, nt in enumerate([64, 48, 64])]
    poisoned = [np.array(x, copy=True) for x in reqs[1]["state"]]
    poisoned[0][10, 7] = np.nan
    with K.Engine() as eng:
        ids = [eng.admit(r["nt"], r["dt"], *(poisoned if k == 1 else r["state"]), steps=30)
               for k, r in enumerate(reqs)]
        eng.step(30)
        eng.sync()
        assert eng.poll(ids[1]).state == K.engine.REQ_FAILED
        assert eng.poll(ids[1]).steps_done <= 17
        for k in (0, 2):
            assert eng.poll(ids[k]).state == K.engine.REQ_DONE
            got = eng.fetch(ids[k], reqs[k]["nt"])
            ref = alone(reqs[k]["nt"], reqs[k]["dt"], reqs[k]["state"], 30)
            for g, x in zip(got, ref):
                np.testing.assert_array_equal(g, x)
        assert eng.stats().failed == 1
        with pytest.raises(K.KaminoError):
            eng.fetch(ids[1], reqs[1]["nt"])
    with K.Sim(64, 0.01) as s:
        st = [np.array(x, copy=True) for x in reqs[0]["state"]]
        s.set_fields(*st)
        s.step(2)
        assert s.status() == 0
        st[2][3, 3] = np.inf
        s.set_fields(*st)
        s.step(1)
        assert s.status() & 1


def test_overflowed_velocities_finish_and_are_flagged():
    """Infinite u_phi and huge u_theta in a band of rows run through every
    kernel (checked sampler, the cubic's FP64 branch, the transforms) in
    bounded time and raise the watchdog word."""
    r = make_request(64, 9)
    st = [np.array(x, copy=True) for x in r["state"]]
    st[0][20:28, :] = np.inf
    st[1][20:28, :] = -1e30
    with K.Sim(64, 0.01) as s:
        s.set_fields(*st)
        s.step(2)
        s.sync()
        assert s.status() & 1
    log("[watchdog] infinite velocities: steps completed and flagged")


def test_autotune_runs_and_tile_shapes_never_change_values(tmp_path):
    r = make_request(96, 5)
    base = alone(96, r["dt"], r["state"], 4)
    cache = str(tmp_path / "tune.txt")
    k = K.autotune(96, batch=4, steps=2, cache_path=cache)
    assert (k > 0).all()
    assert open(cache).read().startswith("gfx950 96 4 ")
    tuned = K.tune_get(96)
    try:
        for shape in [(4, 64, 2), (16, 96, 1), (32, 32, 2), tuned]:
            K.tune_set(96, *shape)
            got = alone(96, r["dt"], r["state"], 4)
            with K.Engine() as eng:
                i = eng.admit(96, r["dt"], *r["state"], steps=4)
                eng.step(4)
                got_e = eng.fetch(i, 96)
            for a, b, c in zip(got, got_e, base):
                np.testing.assert_array_equal(a, c)
                np.testing.assert_array_equal(b, c)
    finally:
        K.tune_set(96, 0, 0, 0)
    log(f"[tune] 96: tuned {tuned}, best kernel ms {np.round(k, 4)}")


def test_cancel_of_a_finished_request_releases_it():
    """Cancelling a request that has finished (its final copy landed or still
    in flight) forgets it: no completion query reports it afterwards, a fetch
    finds nothing, and the engine keeps serving; its pinned copy goes back to
    the pool (a later request of the same size reuses it)."""
    r = make_request(48, 21)
    with K.Engine() as eng:
        a = eng.admit(48, r["dt"], *r["state"], steps=2)
        b = eng.admit(48, r["dt"], *r["state"], steps=2)
        eng.step(2)  # both finish; their copies are queued
        eng.cancel(a)  # may still be in flight
        eng.sync()
        eng.cancel(b)  # landed, listed as completed
        assert eng.completed() == []
        for rid in (a, b):
            with pytest.raises(K.KaminoError):
                eng.fetch(rid, 48)
        c = eng.admit(48, r["dt"], *r["state"], steps=2)
        eng.step(2)
        eng.sync()
        assert eng.completed() == [c]
        got = eng.fetch(c, 48)
        for g, x in zip(got, alone(48, r["dt"], r["state"], 2)):
            np.testing.assert_array_equal(g, x)


def test_4096_slots_rebuild_cost_and_results():
    """The engine at its default slot count: 4096 small spheres admitted, stepped
    together, one cancelled (a membership change rebuilds the descriptor and
    tile tables on the host); the host time of a step with a rebuild is logged
    beside a plain step, and two of the spheres equal their solo runs."""
    import time

    nt, n = 32, 4096
    st = seeded_state(nt, L=8, seed=9, dens=1)
    st2 = seeded_state(nt, L=8, seed=10, dens=1)
    with K.Engine(max_slots=n) as eng:
        ids = [eng.admit(nt, 0.5 / nt, *(st2 if k == 7 else st), steps=6) for k in range(n)]
        assert all(i is not None for i in ids)
        t0 = time.perf_counter()
        eng.step(1)  # first step: builds the batch of 4096
        eng.sync()
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        eng.step(1)
        eng.sync()
        t_plain = time.perf_counter() - t0
        eng.cancel(ids[100])
        t0 = time.perf_counter()
        eng.step(1)  # rebuild for 4095
        eng.sync()
        t_rebuild = time.perf_counter() - t0
        eng.step(3)
        eng.sync()
        got0, got7 = eng.fetch(ids[0], nt), eng.fetch(ids[7], nt)
    log(f"[engine 4096] step with first build {t_build * 1e3:.2f} ms, plain step {t_plain * 1e3:.2f} ms, "
        f"step with rebuild after a cancel {t_rebuild * 1e3:.2f} ms")
    for got, state in ((got0, st), (got7, st2)):
        for a, b in zip(got, alone(nt, 0.5 / nt, state, 6)):
            np.testing.assert_array_equal(a, b)
