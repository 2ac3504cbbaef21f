#!/usr/bin/env python3
"""Benchmark of the spherical-fluid time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.  The workload is B simulations of 256x512 stepped together by the batching engine; a step is one full time step
(advection + geometric + FFT/tridiagonal projection) of all B spheres.  The
plain form (one sphere, its own kernel sequence, no graph) is measured beside
it.  Inputs are resident in HBM before the timed region; timing uses HIP
events on the stream the kernels run on (inside libkamino.so).
"""
import argparse
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9  # MI355X: 256 CUs, 64 VALU lanes per CU per clock, 2.4 GHz
# algorithmic bytes per cell per launch (FP32 compulsory reads + writes)
BYTES_PER_CELL = {"advect": 24.0, "div_fft": 12.0, "tridiag": 8.0, "ifft_grad": 20.0}
KERNELS = ["advect", "div_fft", "tridiag", "ifft_grad"]
METRIC = "cell-steps/s at 256x512 (plain vs batched/served) + p99 step latency; HBM GB/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed batched steps")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=128, help="spheres in the batched workload")
    ap.add_argument("--n-theta", type=int, default=256)
    ap.add_argument("--plain-steps", type=int, default=1000)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uneven-steps", type=int, default=200, help="steps of the uneven 64-sphere batch (0 skips)")
    ap.add_argument("--large-steps", type=int, default=500, help="steps of the one 1024x2048 sphere (0 skips)")
    ap.add_argument("--large-transform", choices=["auto", "fft", "dft"], default="auto",
                    help="phi transform of the 1024x2048 leg: the autotuner's pick, or forced")
    ap.add_argument("--serve-requests", type=int, default=10000, help="requests of the served leg (0 skips)")
    ap.add_argument("--launch", choices=["graph", "direct"], default="graph",
                    help="engine steps of the batched legs as hipGraph replays or as direct kernel "
                         "launches (the uneven-batch leg measures both; the served leg launches directly)")
    ap.add_argument("--graph-steps", type=int, default=32,
                    help="steps captured per hipGraph (each graph launch boundary idles the GPU ~10 us)")
    ap.add_argument("--groups", type=int, default=3,
                    help="concurrent engine groups, each on a stream of its own (kernels of different "
                         "groups overlap)")
    ap.add_argument("--tune-cache", default=None,
                    help="tuning cache: sizes it holds are loaded, the others autotuned and appended")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (for HBM counter scaling)")
    ap.add_argument("--clock-warmup-s", type=float, default=CLOCK_WARMUP_S,
                    help="untimed steps for this many seconds before each leg's warm-up steps "
                         "(the GPU's clocks ramp over the first ~50 ms of work)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


CLOCK_WARMUP_S = 0.3  # --clock-warmup-s


def clock_warmup_engine(eng, seconds, chunk=32):
    """Runs untimed engine steps for `seconds` of wall time before a leg's own
    warm-up steps: the GPU raises its clocks over the first ~50 ms of steady
    work (measured on the 128-sphere batch: 5 warm-up steps give 0.39-0.40 ms
    per step, 200 give 0.331), so a short warm-up would time the ramp, not the
    step.  Steps go out `chunk` at a time (the graph length of a graph-replayed
    leg, so its graphs are captured and first launched here).  Returns the
    number of steps run."""
    import time

    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        eng.step(chunk)
        eng.sync()
        n += chunk
    return n


def clock_warmup_sim(sim, seconds):
    import time

    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        sim.benchmark(50)
        n += 50
    return n


def init_state(nt, L, seed, dens):
    import kaminogpu_amd as K

    return list(K.init_fields(nt, L, seed, dens))


def cpu_baseline(nt, dt, seconds):
    """FP64 oracle on all host cores, a bounded sample of the same workload."""
    from oracle import oracle as O

    up, ut, rho = O.init_fields(nt, 2 * nt, 16, 1, 1)
    threads = O.max_threads()
    t = O.run(nt, 2 * nt, dt, up, ut, rho, 2, threads)
    steps = int(max(2, min(2000, seconds / max(t / 2, 1e-6))))
    t = O.run(nt, 2 * nt, dt, up, ut, rho, steps, threads)
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in txt.splitlines() if l.startswith("Model name")), "")
    except Exception:
        model = ""
    return {
        "value": steps * nt * 2 * nt / t, "unit": "cell-steps/s", "cores": threads, "kind": "port",
        "sample": f"{steps} FP64 steps of one {nt}x{2 * nt} sphere (dt {dt}) on {threads} OpenMP "
                  f"threads, {model}".strip(),
    }


TUNED = set()
USE_GRAPH = False  # set from --launch
GRAPH_STEPS = 32  # set from --graph-steps
GROUPS = 3  # set from --groups


def tune(K, nt, batch, cache):
    """Tile shapes for n_theta: from the cache when it holds them, else timed on
    the device (kam_autotune) and appended to the cache.  Returns the tiles and
    the best per-kernel times (None when loaded)."""
    if cache and os.path.exists(cache) and not TUNED:
        TUNED.update(K.tune_load(cache))
    ms = None
    if nt not in TUNED:
        ms = K.autotune(nt, batch, steps=10, cache_path=cache)
        TUNED.add(nt)
    tiles = dict(zip(["adv_rows", "adv_cols", "fft_rows"], K.tune_get(nt)))
    # the phi transform the autotuner picked for this size: the row FFT or the MFMA DFT
    tiles["transform"] = {K.TRANSFORM_DFT: "dft_mfma", K.TRANSFORM_FFT8: "fft_w8"}.get(K.tune_get_transform(nt), "fft")
    return tiles, ms


def run_uneven(K, steps, cache, sims=64, seed=3):
    """The uneven batch: 64 spheres with n_theta drawn from {32,...,256}, dt = 0.5/n_theta,
    L = min(16, n_theta/4), density band, stepped together through one engine;
    tiles of each size from the autotuner.  Steps are replayed from hipGraphs of
    GRAPH_STEPS steps each in GROUPS concurrent groups; the same leg with direct
    launches in one group is reported beside it, and the per-kernel times are
    those of one group (one launch per kernel covering the whole batch)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([32, 48, 64, 96, 128, 192, 256], sims)
    tiles = {}
    for nt in sorted(set(int(n) for n in sizes)):
        # timed on at least 64 spheres of the size: a handful of small spheres
        # is launch latency, and its timings pick tiles at random
        tiles[nt], _ = tune(K, nt, max(64, int((sizes == nt).sum())), cache)
    log(f"uneven tiles: {tiles}")
    forms = {}
    for form in ("graph", "direct"):
        graph = form == "graph"
        with K.Engine(use_graph=graph, graph_steps=GRAPH_STEPS, groups=GROUPS if graph else 1) as eng:
            for k, nt in enumerate(sizes):
                st = K.init_fields(int(nt), min(16, int(nt) // 4), 100 + k, K.DENSITY_BAND)
                eng.admit(int(nt), 0.5 / nt, *st, steps=10 ** 6)
            clock_warmup_engine(eng, CLOCK_WARMUP_S, GRAPH_STEPS)
            eng.step(16)
            eng.reset_stats()
            eng.step(steps)
            s = eng.stats()
            if not graph:
                kern = eng.profile(10)
        v = s.cell_steps / (s.gpu_ms * 1e-3)
        forms[form] = {"cell_steps_per_s": v, "ms_per_step": s.gpu_ms / s.steps,
                       "p50_step_ms": s.step_p50_ms, "p99_step_ms": s.step_p99_ms}
        log(f"uneven x{sims} ({form}): {v:.3e} cell-steps/s, {s.gpu_ms / s.steps:.4f} ms/step, "
            f"p50 {s.step_p50_ms:.4f} p99 {s.step_p99_ms:.4f} ms")
    log(f"uneven kernel ms {np.round(kern, 4)}")
    g = forms["graph"]
    return {"sims": sims, "cells": int(sum(2 * int(n) ** 2 for n in sizes)), "steps": steps,
            "launch": f"hipGraph of {GRAPH_STEPS} steps", "groups": GROUPS, "cell_steps_per_s": g["cell_steps_per_s"],
            "p50_step_ms": g["p50_step_ms"], "p99_step_ms": g["p99_step_ms"], "direct_launches": forms["direct"],
            "kernel_ms": dict(zip(["advect", "div_fft", "tridiag", "ifft_grad"], kern.tolist())),
            "tiles": {str(k): t for k, t in tiles.items()}}


def run_served(K, n):
    """Serving: Poisson arrivals at 80% of the mix's saturated capacity (128
    requests of the same distribution stepped together); n_theta log-uniform
    in [32, 512] (rounded to supported sizes), steps U{50..2000}."""
    from kaminogpu_amd import serve

    reqs = serve.make_requests(n, seed=4)
    work = float(np.mean([2 * r.n_theta ** 2 * r.steps for r in reqs]))
    pool = serve.StatePool()
    # direct launches: membership changes every few tens of steps, and each
    # change would re-capture the graphs
    with K.Engine(groups=GROUPS) as eng:
        cap = serve.capacity(serve.make_requests(128, seed=5), eng, pool)
        rate = 0.8 * cap / work
        log(f"serve: mix capacity {cap:.3e} cell-steps/s -> {rate:.1f} requests/s")
        out = serve.serve(reqs, rate, eng, pool, log=log)
    out["load"] = out["cell_steps_per_s"] / cap
    log(f"served {n}: {out['cell_steps_per_s']:.3e} cell-steps/s ({out['load']:.2f} of capacity), "
        f"GPU in steps {out['gpu_busy']:.2f} of the wall time, "
        f"step p50 {out['step_p50_ms']:.3f} p99 {out['step_p99_ms']:.3f} ms, request p50 "
        f"{out['request_p50_s']:.3f} p99 {out['request_p99_s']:.3f} s")
    out["mean_cell_steps_per_request"] = work
    out["mix_capacity_cell_steps_per_s"] = cap
    return out


def run_large(K, steps, cache, nt=1024, transform="auto"):
    """One large sphere: 1024x2048, dt = 0.5/nTheta, seeded flow with L = 32,
    stepped alone (plain launches, FP32 fields and pressure); tiles autotuned
    on the sphere itself.  transform "fft" or "dft" overrides the autotuner's
    choice of the phi transform (the MFMA DFT products, for profiling them).
    The same sphere as the only member of an engine, replayed from hipGraphs of
    GRAPH_STEPS steps, is reported beside it."""
    dt = 0.5 / nt
    tiles, _ = tune(K, nt, 1, cache)
    if transform != "auto":
        K.tune_set_transform(nt, K.TRANSFORM_DFT if transform == "dft" else K.TRANSFORM_FFT)
        tiles["transform"] = "dft_mfma" if transform == "dft" else "fft"
    log(f"large tiles: {tiles}")
    with K.Sim(nt, dt) as s:
        s.set_fields(*init_state(nt, 32, 2, 1))
        clock_warmup_sim(s, CLOCK_WARMUP_S)
        s.benchmark(10)
        total, per, _ = s.benchmark(steps)
        _, _, kern = s.benchmark(10, per_kernel=True)
    v = nt * 2 * nt * steps / (total * 1e-3)
    log(f"large {nt}x{2 * nt}: {v:.3e} cell-steps/s, {total / steps:.4f} ms/step, kernel ms {np.round(kern, 4)}")
    with K.Engine(use_graph=True, graph_steps=GRAPH_STEPS) as eng:
        eng.admit(nt, dt, *init_state(nt, 32, 2, 1), steps=10 ** 7)
        clock_warmup_engine(eng, CLOCK_WARMUP_S, GRAPH_STEPS)
        eng.reset_stats()
        eng.step(steps)
        sg = eng.stats()
    vg = sg.cell_steps / (sg.gpu_ms * 1e-3)
    log(f"large {nt}x{2 * nt} in an engine, {GRAPH_STEPS}-step graphs: {vg:.3e} cell-steps/s, "
        f"{sg.gpu_ms / sg.steps:.4f} ms/step, p99 {sg.step_p99_ms:.4f}")
    return {"n_theta": nt, "steps": steps, "cell_steps_per_s": v, "ms_per_step": total / steps,
            "p99_step_ms": float(np.percentile(per, 99)), "tiles": tiles,
            "kernel_ms": dict(zip(KERNELS, [float(x) for x in kern])),
            "engine_graph": {"launch": f"hipGraph of {GRAPH_STEPS} steps", "cell_steps_per_s": vg,
                             "ms_per_step": sg.gpu_ms / sg.steps, "p99_step_ms": sg.step_p99_ms}}


def profile_counters(kernel, nt, batch):
    """Per-launch counters of `kernel` from the committed rocprofv3 summary of
    this workload: HBM bytes (FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel
    calibration) and VALU instructions; empty when the summary is for another
    workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{batch}x{nt}x{2 * nt}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        t = json.load(f)
    name = {"advect": "k_advect", "div_fft": "k_div_fft", "tridiag": "k_tridiag",
            "ifft_grad": "k_ifft_grad"}[kernel]
    return t["kernels"].get(name) or {}


def copy_calibration(nbytes):
    """One device fill and one device-to-device copy of nbytes through the HIP
    runtime: kernels of known traffic that scale the HBM byte counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes filled, then copied device to device")


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Streaming bandwidth of this GPU as HIP measures it: device-to-device
    copies of nbytes (read + write), timed with HIP events after one warm-up."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    assert hip.hipMemset(a, 1, n) == 0
    assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice, warm-up
    assert hip.hipEventRecord(e0, None) == 0
    for _ in range(reps):
        assert hip.hipMemcpyAsync(b, a, n, 3, None) == 0
    assert hip.hipEventRecord(e1, None) == 0 and hip.hipEventSynchronize(e1) == 0
    ms = C.c_float()
    assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
    for ev in (e0, e1):
        hip.hipEventDestroy(ev)
    hip.hipFree(a)
    hip.hipFree(b)
    gbs = 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    log(f"streaming copy: {gbs:.0f} GB/s ({reps} device-to-device copies of {nbytes >> 20} MiB)")
    return gbs


def main():
    a = parse()
    global USE_GRAPH, GRAPH_STEPS, GROUPS, CLOCK_WARMUP_S
    USE_GRAPH = a.launch == "graph"
    GRAPH_STEPS = a.graph_steps
    GROUPS = a.groups
    CLOCK_WARMUP_S = a.clock_warmup_s
    if a.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X; "
              "multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)
    import kaminogpu_amd as K

    if K.device_count() < 1:
        raise SystemExit("no gfx950 device visible: " + K.lib().kam_last_error().decode())
    nt, dt, B = a.n_theta, 0.002, a.batch
    cells = nt * 2 * nt
    if a.copy_calibration_mib > 0:
        copy_calibration(a.copy_calibration_mib << 20)
    copy_gbs = copy_bandwidth()

    # ---- plain form: one simulation, per-step kernel sequence, no graph
    plain, p_steps = None, np.zeros(1)
    if a.plain_steps > 0:
        state = init_state(nt, 16, 1, 1)
        with K.Sim(nt, dt) as s:
            s.set_fields(*state)
            clock_warmup_sim(s, CLOCK_WARMUP_S)
            s.benchmark(20)
            p_total, p_steps, _ = s.benchmark(a.plain_steps)
        plain = cells * a.plain_steps / (p_total * 1e-3)
        log(f"plain: {plain:.3e} cell-steps/s, {p_total / a.plain_steps:.4f} ms/step")

    # ---- tile shapes for this sphere size, timed on the device (kam_autotune)
    tiles, tuned_ms = tune(K, nt, B, a.tune_cache)
    log(f"tiles {nt}: {tiles}" + ("" if tuned_ms is None else f", autotuned, kernel ms {np.round(tuned_ms, 4)}"))

    # ---- batched form: B spheres through the engine (graph-replayed or directly launched steps)
    states = [init_state(nt, 16, 1 + b, 1) for b in range(min(B, 8))]

    def batched_leg(precision):
        def admitted(groups):
            eng = K.Engine(use_graph=USE_GRAPH, graph_steps=GRAPH_STEPS, groups=groups)
            for b in range(B):
                if eng.admit(nt, dt, *states[b % len(states)], steps=10 ** 6, precision=precision) is None:
                    raise SystemExit("engine refused the batch")
            return eng
        with admitted(GROUPS) as eng:
            clock_steps = clock_warmup_engine(eng, CLOCK_WARMUP_S, GRAPH_STEPS)
            eng.step(a.warmup)
            eng.reset_stats()
            eng.step(a.steps)
            st = eng.stats()
        # per-kernel times of the whole batch as one launch per kernel (one
        # group, HIP events around each launch): the roofline's kernel, as the
        # rocprof kernel trace of `--groups 1` records it
        with admitted(1) as eng:
            clock2 = clock_warmup_engine(eng, 0.1, GRAPH_STEPS)
            kern_ms = eng.profile(a.profile_steps)
        return st, kern_ms, (clock_steps, clock2)

    st, kern_ms, (clock_steps, clock2) = batched_leg(K.PREC_FP32)
    ms_per_step = st.gpu_ms / st.steps
    batched = st.cell_steps / (st.gpu_ms * 1e-3)
    log(f"batched x{B}: {batched:.3e} cell-steps/s, {ms_per_step:.4f} ms/step, "
        f"p50 {st.step_p50_ms:.4f} p99 {st.step_p99_ms:.4f}")
    # the same batch with the FP64 pressure (post-projection divergence at the
    # FP32 storage floor)
    st64, kern64, _ = batched_leg(K.PREC_FP64_PRESSURE)
    p64 = {"cell_steps_per_s": st64.cell_steps / (st64.gpu_ms * 1e-3), "ms_per_step": st64.gpu_ms / st64.steps,
           "p50_step_ms": st64.step_p50_ms, "p99_step_ms": st64.step_p99_ms,
           "kernel_ms": dict(zip(KERNELS, [float(x) for x in kern64]))}
    log(f"batched x{B} FP64 pressure: {p64['cell_steps_per_s']:.3e} cell-steps/s, "
        f"{p64['ms_per_step']:.4f} ms/step, kernel ms {np.round(kern64, 4)}")

    large = run_large(K, a.large_steps, a.tune_cache, transform=a.large_transform) if a.large_steps > 0 else None
    uneven = run_uneven(K, a.uneven_steps, a.tune_cache) if a.uneven_steps > 0 else None
    served = run_served(K, a.serve_requests) if a.serve_requests > 0 else None

    kern = {k: float(v) for k, v in zip(KERNELS, kern_ms)}
    dom = max(kern, key=kern.get)
    gbs = {k: BYTES_PER_CELL[k] * cells * B / (kern[k] * 1e-3) / 1e9 for k in KERNELS}
    achieved = gbs[dom]
    counters = profile_counters(dom, nt, B)
    traffic = counters.get("hbm_bytes")
    # the same kernel against the VALU issue peak: wave instructions x 64 lanes
    # per launch over the launch time, against 256 CUs x 64 lanes x 2.4 GHz
    valu = counters.get("valu_insts")
    valu_frac = None if valu is None else valu * 64 / (kern[dom] * 1e-3) / VALU_PEAK_LANE_OPS
    out = {
        "metric": METRIC,
        "value": batched,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded spherical-harmonic streamfunction, density band)",
        "config": {"workload": f"batched: {B} spheres of {nt}x{2 * nt}, dt {dt}, one engine step "
                               f"advances all {B}", "n_theta": nt, "n_phi": 2 * nt, "sims": B,
                   "tiles": tiles, "launch": a.launch if a.launch == "direct" else
                   f"hipGraph of {a.graph_steps} steps", "groups": a.groups, "clock_warmup_steps": clock_steps,
                   "profile_clock_warmup_steps": clock2},
        "batched": {"cell_steps_per_s": batched, "p50_step_ms": st.step_p50_ms,
                    "p99_step_ms": st.step_p99_ms, "speedup_vs_plain": batched / plain if plain else None},
        "batched_fp64_pressure": p64,
        "plain": {"cell_steps_per_s": plain, "steps": a.plain_steps,
                  "p50_step_ms": float(np.percentile(p_steps, 50)),
                  "p99_step_ms": float(np.percentile(p_steps, 99))},
        "large": large,
        "uneven_batch": uneven,
        "served": served,
        "kernel_ms": kern,
        "kernel_gbs": gbs,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "copy_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "algorithmic_bytes": BYTES_PER_CELL[dom] * cells * B,
                     "valu_issue_frac": valu_frac},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nt, dt, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
