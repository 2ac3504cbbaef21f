"""Serving driver: a Poisson-arrival stream of simulation requests through the
engine (continuous batching).

Requests arrive on the wall clock; each waits in a FIFO queue until the
engine admits it (a free slot and room in the HBM arena), then runs its own
number of steps inside the engine's shared batched steps, and retires when its
final fields reach the host.  Reported: cell-steps/s over the run, per-step
latency percentiles (GPU time of each shared step: from the previous step's
end, or from its own start when the GPU had gone idle), the fraction of the
wall time the GPU spent in steps, and per-request latency percentiles (arrival
to results on the host).
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import DENSITY_NOISE, init_fields
from .engine import Engine


def smooth_sizes(lo=32, hi=512):
    """n_theta values the row FFT supports (2^a 3^b 5^c) in [lo, hi]."""
    out = []
    for n in range(lo, hi + 1):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return np.array(out)


@dataclass
class Request:
    n_theta: int
    steps: int
    seed: int
    arrive: float = 0.0
    admit: float = math.nan
    done: float = math.nan


def make_requests(n, seed=0, lo=32, hi=512, min_steps=50, max_steps=2000):
    """n_theta log-uniform in [lo, hi] rounded to the nearest supported size;
    steps uniform in [min_steps, max_steps]."""
    rng = np.random.default_rng(seed)
    sizes = smooth_sizes(lo, hi)
    raw = np.exp(rng.uniform(math.log(lo), math.log(hi), n))
    nts = sizes[np.abs(np.log(sizes)[None, :] - np.log(raw)[:, None]).argmin(axis=1)]
    steps = rng.integers(min_steps, max_steps + 1, n)
    return [Request(int(a), int(b), int(seed * 1_000_003 + k)) for k, (a, b) in enumerate(zip(nts, steps))]


class StatePool:
    """Initial states by (n_theta, seed % variants): the seeded generator is run
    once per pool entry, so request setup stays off the serving loop's path."""

    def __init__(self, L=8, variants=4):
        self.L, self.variants, self.cache = L, variants, {}

    def get(self, n_theta, seed):
        key = (n_theta, seed % self.variants)
        if key not in self.cache:
            self.cache[key] = init_fields(n_theta, self.L, key[1] + 1, DENSITY_NOISE)
        return self.cache[key]


def capacity(requests, engine: Engine, pool: StatePool | None = None, steps=50, dt_scale=0.5):
    """Saturated throughput of a request mix: every request is admitted at once
    (as far as the engine's budget allows) and stepped together; returns
    cell-steps/s of GPU step time.  The requests are cancelled afterwards."""
    pool = pool or StatePool()
    ids = []
    for r in requests:
        rid = engine.admit(r.n_theta, dt_scale / r.n_theta, *pool.get(r.n_theta, r.seed),
                           steps=steps + 10)
        if rid is None:
            break
        ids.append(rid)
    engine.step(5)
    engine.reset_stats()
    engine.step(steps)
    st = engine.stats()
    for rid in ids:
        engine.cancel(rid)
    engine.sync()
    return st.cell_steps / (st.gpu_ms * 1e-3)


def serve(requests, rate, engine: Engine, pool: StatePool | None = None, dt_scale=0.5,
          arrival_seed=1, log=None):
    """Runs the stream; `rate` in requests/s.  dt = dt_scale / n_theta."""
    pool = pool or StatePool()
    for r in requests:  # warm the state pool before the clock starts
        pool.get(r.n_theta, r.seed)
    rng = np.random.default_rng(arrival_seed)
    gaps = rng.exponential(1.0 / rate, len(requests))
    arrivals = np.cumsum(gaps) - gaps[0]
    queue = deque()
    running = {}
    nxt = done = 0
    engine.reset_stats()
    t0 = engine.now()
    last_log = t0
    while done < len(requests):
        now = engine.now() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            requests[nxt].arrive = arrivals[nxt]
            queue.append(requests[nxt])
            nxt += 1
        while queue:
            r = queue[0]
            rid = engine.admit(r.n_theta, dt_scale / r.n_theta, *pool.get(r.n_theta, r.seed),
                               steps=r.steps)
            if rid is None:
                break
            r.admit = engine.now() - t0
            running[rid] = r
            queue.popleft()
        if running:
            engine.step(1)
        elif nxt < len(requests):
            time.sleep(max(0.0, min(1e-3, arrivals[nxt] - now)))
        for rid in engine.completed():
            r = running.pop(rid)
            r.done = engine.poll(rid).done_s - t0
            engine.fetch_release(rid)
            done += 1
        if log and engine.now() - last_log > 5.0:
            last_log = engine.now()
            log(f"serve: t={last_log - t0:.1f}s done {done}/{len(requests)} running {len(running)} "
                f"queued {len(queue)}")
    engine.sync()
    st = engine.stats()
    wall = engine.now() - t0
    lat = np.array([r.done - r.arrive for r in requests])
    wait = np.array([r.admit - r.arrive for r in requests])
    return {
        "requests": len(requests),
        "rate_per_s": rate,
        "wall_s": wall,
        "engine_steps": int(st.steps),
        "cell_steps": int(st.cell_steps),
        "cell_steps_per_s": st.cell_steps / wall,
        "step_p50_ms": st.step_p50_ms,
        "step_p99_ms": st.step_p99_ms,
        "gpu_busy": st.gpu_ms * 1e-3 / wall,
        "request_p50_s": float(np.percentile(lat, 50)),
        "request_p99_s": float(np.percentile(lat, 99)),
        "queue_wait_p99_s": float(np.percentile(wait, 99)),
    }
