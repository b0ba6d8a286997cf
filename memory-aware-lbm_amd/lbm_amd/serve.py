"""Continuous-batching solve server for cylinder-channel requests on one GPU.

A request is one complete solve: a channel (size, disc, Re), a step count and
a storage format (FP64 bit-exact, or FP32 deviations).  Requests arrive on
their own clock (a seeded Poisson stream for the serving workload).  The
server loop:

  * admits waiting requests first-fit, in arrival order, as long as a slot is
    free (max_active) and the device-memory footprint of the new lattice
    (with the further buffers of the sweep it will run) fits the HBM budget
    next to the solves already resident;
  * keeps every admitted lattice resident on the device until its solve ends;
  * hands each resident solve a time slice of steps (about slice_ms of
    work, a multiple of the captured graph length) on the solve's own HIP
    stream whenever its previous slice has finished, so solves of different
    sizes run concurrently on the GPU and a newly admitted request starts at
    the next slice boundary instead of waiting for the queue to drain; each
    solve runs the sweep fastest for its size and format (SolveServer.SWEEPS);
  * on completion takes the solve's diagnostics (and, if asked, its velocity
    field), frees its memory and records arrival -> completion latency.

Contexts allocate from a stream-ordered device memory pool the server
reserves at start (lbm_hip_pool_reserve), and every call on a context waits
only for that context's stream, so admitting or completing a solve never
stalls the solves already running.

A solve's result does not depend on what else is resident: every solve runs
in its own device context, on the same kernels, so its numbers are the same
alone and in the mix (bit for bit; tests/test_gpu_serve.py checks it).

Workload of the serving configuration: lx log-uniform in [256, 4096],
ly = lx/4, disc as in the batching workload, Re uniform in [20, 500], steps
uniform in [1e3, 5e4], storage FP64 or FP32 with equal probability.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional

import numpy as np

from . import Channel, Lattice, device_memory, footprint_of, pool_release, pool_reserve


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream starts
    channel: Channel
    steps: int
    storage: str = "f64"


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    cells: int
    storage: str
    diagnostics: np.ndarray
    fields: Optional[tuple] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def queued(self) -> float:
        return self.admitted - self.arrival


def poisson_requests(n: int, rate: float, seed: int = 0, lx_lo: int = 256,
                     lx_hi: int = 4096, re_lo: float = 20.0, re_hi: float = 500.0,
                     steps_lo: int = 1000, steps_hi: int = 50000,
                     storages=("f64", "f32")) -> List[Request]:
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(rng.exponential(1.0 / rate))
        lx = int(np.exp(rng.uniform(np.log(lx_lo), np.log(lx_hi))))
        ly = max(lx // 4, 3)
        ox = int(rng.uniform(0.15, 0.3) * lx)
        oy = int(rng.uniform(0.4, 0.6) * ly)
        r = max(int(rng.uniform(0.05, 0.12) * ly), 1)
        re = float(rng.uniform(re_lo, re_hi))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        storage = storages[int(rng.integers(0, len(storages)))]
        out.append(Request(rid, t, Channel(lx, ly, ox, oy, r, 0.02, re), steps, storage))
    return out


@dataclass
class _Active:
    req: Request
    lat: Lattice
    admitted: float
    remaining: int
    slice_steps: int
    footprint: int
    inflight: bool = False


class SolveServer:
    """Admission by device-memory footprint, device-resident lattices,
    time-sliced concurrent execution (one HIP stream per solve)."""

    # variant="auto" picks the sweep that is fastest for one solve of that
    # size and format: SWEEPS lists (minimum cells, variant) in increasing
    # size.  Measured on MI355X, one solve alone, graphs of 50 steps, GLUPS
    # fused / tight_block / panel / panel4 / panel6:
    #   FP64 512 x 128 9.8 / 9.8 / 4.7 / 4.7 / 4.7, 768 x 192 18.4 / 14.2 /
    #   10.6 / 8.3 / 8.7, 1024 x 256 22.8 / 23.8 / 17.8 / 14.4 / 15.1,
    #   1536 x 384 26.7 / 36.5 / 31.2 / 31.4 / 32.3, 2048 x 512 33.8 / 43.1 /
    #   44.7 / 46.3 / 51.0, 3072 x 768 38.3 / 53.0 / 50.4 / 69.4 / 76.2,
    #   4096 x 1024 38.1 / 57.7 / 52.9 / 82.6 / 95.8, 8192 x 2048 39.2 / 58.9 /
    #   61.1 / 109.8 / 134.0;
    #   FP32 512 x 128 8.1 / 5.4 / 4.2 / 4.2 / 4.2, 1024 x 256 26.8 / 16.4 /
    #   15.9 / 12.8 / 12.8, 1536 x 384 33.7 / 32.0 / 34.4 / 27.7 / 28.7,
    #   2048 x 512 47.0 / 42.1 / 51.9 / 42.3 / 40.0, 4096 x 1024 68.2 / 73.1 /
    #   90.9 / 85.6 / 79.0, 8192 x 2048 77.3 / 88.1 / 110.7 / 133.1 / 126.6.
    SWEEPS = {"f64": ((0, "combine"), (200_000, "tight_block"), (1_000_000, "panel6")),
              "f32": ((0, "combine"), (500_000, "panel"), (8_000_000, "panel4")),
              "f16": ((0, "combine"), (800_000, "panel"), (8_000_000, "panel4"))}

    def __init__(self, hbm_budget: Optional[int] = None, max_active: int = 8,
                 slice_ms: float = 20.0, graph_steps: int = 50,
                 est_glups: Dict[str, float] = None, keep_fields: bool = False,
                 variant: str = "auto", pool_bytes: Optional[int] = 8 << 30):
        if hbm_budget is None:
            free, _ = device_memory()
            hbm_budget = int(0.9 * free)
        self.hbm_budget = hbm_budget
        # device memory the context pool takes up front (None: no pool, every
        # context uses hipMalloc / hipFree)
        self.pool_bytes = None if pool_bytes is None else min(int(pool_bytes), hbm_budget)
        self.max_active = max_active
        self.slice_ms = slice_ms
        self.graph_steps = max(1, graph_steps)
        self.est_glups = est_glups or {"f64": 35.0, "f32": 70.0, "f16": 110.0}
        self.keep_fields = keep_fields
        self.variant = variant
        self.stats = {"admitted": 0, "peak_active": 0, "peak_bytes": 0, "slices": 0}

    def _variant(self, req: Request) -> str:
        if self.variant != "auto":
            return self.variant
        cells = req.channel.lx * req.channel.ly
        return [v for lo, v in self.SWEEPS[req.storage] if cells >= lo][-1]

    def _slice(self, req: Request) -> int:
        cells = req.channel.lx * req.channel.ly
        per_step_ms = cells / (self.est_glups[req.storage] * 1e9) * 1e3
        n = int(self.slice_ms / max(per_step_ms, 1e-9))
        n = max(self.graph_steps, n // self.graph_steps * self.graph_steps)
        return n

    def run(self, requests: List[Request], clock=time.perf_counter) -> List[Result]:
        if self.pool_bytes is not None:
            pool_reserve(self.pool_bytes)
        try:
            return self._run(requests, clock)
        finally:
            if self.pool_bytes is not None:
                pool_release()

    def _run(self, requests: List[Request], clock) -> List[Result]:
        pending: Deque[Request] = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: List[Request] = []
        active: List[_Active] = []
        results: List[Result] = []
        used = 0
        t0 = clock()
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: first fit in arrival order
            for req in list(waiting):
                if len(active) >= self.max_active:
                    break
                fp = footprint_of(req.channel, req.storage, self._variant(req))
                if fp > self.hbm_budget:
                    raise ValueError(f"request {req.rid} needs {fp} B, budget {self.hbm_budget} B")
                if used + fp > self.hbm_budget:
                    continue
                lat = Lattice(req.channel, storage=req.storage)
                lat.set_graph_steps(self.graph_steps)
                lat.init_equilibrium()
                active.append(_Active(req, lat, clock() - t0, req.steps, self._slice(req), fp))
                waiting.remove(req)
                used += fp
                self.stats["admitted"] += 1
                self.stats["peak_active"] = max(self.stats["peak_active"], len(active))
                self.stats["peak_bytes"] = max(self.stats["peak_bytes"], used)
            # dispatch / completion
            progressed = False
            for a in list(active):
                if a.inflight and not a.lat.done():
                    continue
                a.inflight = False
                if a.remaining > 0:
                    n = min(a.slice_steps, a.remaining)
                    a.lat.step(n, self._variant(a.req))
                    a.remaining -= n
                    a.inflight = True
                    self.stats["slices"] += 1
                    progressed = True
                    continue
                diag = a.lat.diagnostics()
                fields = a.lat.macros() if self.keep_fields else None
                finished = clock() - t0
                results.append(Result(a.req.rid, a.req.arrival, a.admitted, finished,
                                      a.req.steps, a.req.channel.lx * a.req.channel.ly,
                                      a.req.storage, diag, fields))
                a.lat.close()
                used -= a.footprint
                active.remove(a)
                progressed = True
            if not progressed:
                if not active and pending and not waiting:
                    time.sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.05)))
                else:
                    time.sleep(0.0002)
        return sorted(results, key=lambda r: r.rid)


def summarize(results: List[Result], makespan: Optional[float] = None) -> dict:
    lat = np.array([r.latency for r in results])
    q = np.array([r.queued for r in results])
    end = max(r.finished for r in results)
    start = min(r.arrival for r in results)
    span = makespan if makespan is not None else end - start
    lups = sum(r.cells * r.steps for r in results)
    return {"solves": len(results), "solves_per_s": len(results) / span,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "max_latency_s": float(lat.max()),
            "p50_queue_s": float(np.percentile(q, 50)), "p99_queue_s": float(np.percentile(q, 99)),
            "glups_delivered": lups / span / 1e9, "makespan_s": span}
