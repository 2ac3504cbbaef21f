#!/usr/bin/env python3
"""Benchmark of the spherical-fluid time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.  A step is one full time step (advection + geometric +
FFT/tridiagonal projection) of the workload.  Inputs are resident in HBM
before the timed region; timing uses HIP events on the stream the kernels run
on (inside libkamino.so).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak
# algorithmic bytes per cell per launch (FP32 compulsory reads + writes)
BYTES_PER_CELL = {"advect": 24.0, "div_fft": 12.0, "tridiag": 8.0, "ifft_grad": 20.0}
KERNELS = ["advect", "div_fft", "tridiag", "ifft_grad"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n-theta", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def init_state(nt, L, seed, dens):
    from oracle import oracle as O

    up, ut, rho = O.init_fields(nt, 2 * nt, L, seed, dens)
    return [np.ascontiguousarray(a, np.float32) for a in (up, ut, rho)]


def cpu_baseline(nt, dt, seconds):
    """FP64 oracle on all host cores, a bounded sample of the same workload."""
    from oracle import oracle as O

    up, ut, rho = O.init_fields(nt, 2 * nt, 16, 1, 1)
    threads = O.max_threads()
    t = O.run(nt, 2 * nt, dt, up, ut, rho, 2, threads)
    steps = int(max(2, min(2000, seconds / max(t / 2, 1e-6))))
    t = O.run(nt, 2 * nt, dt, up, ut, rho, steps, threads)
    try:
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in model.splitlines() if l.startswith("Model name")), "")
    except Exception:
        model = ""
    return {
        "value": steps * nt * 2 * nt / t, "unit": "cell-steps/s", "cores": threads, "kind": "port",
        "sample": f"{steps} FP64 steps of one {nt}x{2 * nt} sphere (dt {dt}), OpenMP, {model}".strip(),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has one MI355X; "
              "multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)
    import kaminogpu_amd as K

    if K.device_count() < 1:
        raise SystemExit("no gfx950 device visible: " + K.lib().kam_last_error().decode())
    nt, dt = a.n_theta, 0.002
    cells = nt * 2 * nt
    state = init_state(nt, 16, 1, 1)

    # plain form: one simulation, per-step kernel sequence, no graph
    with K.Sim(nt, dt) as s:
        s.set_fields(*state)
        s.benchmark(max(1, a.warmup))
        total_ms, step_ms, _ = s.benchmark(a.steps)
        _, _, kern_ms = s.benchmark(min(a.steps, 200), per_kernel=True)
    plain = cells * a.steps / (total_ms * 1e-3)

    kern = {k: float(v) for k, v in zip(KERNELS, kern_ms)}
    dom = max(kern, key=kern.get)
    achieved = BYTES_PER_CELL[dom] * cells / (kern[dom] * 1e-3) / 1e9
    out = {
        "metric": "cell-steps/s at 256x512 (plain vs batched/served) + p99 step latency; HBM GB/s",
        "value": plain,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": total_ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded spherical-harmonic streamfunction, density band)",
        "config": {"workload": f"plain: one {nt}x{2 * nt} sphere, dt {dt}", "n_theta": nt,
                   "n_phi": 2 * nt, "sims": 1},
        "plain": {"cell_steps_per_s": plain, "p50_step_ms": float(np.percentile(step_ms, 50)),
                  "p99_step_ms": float(np.percentile(step_ms, 99))},
        "kernel_ms": kern,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nt, dt, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
