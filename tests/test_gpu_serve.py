"""The solve server: results are identical alone and in the mix, admission
respects the HBM budget and the slot count, late arrivals are admitted while
earlier solves are still running."""
import numpy as np
import pytest

import lbm_amd as L
from lbm_amd.serve import Request, SolveServer, poisson_requests, summarize

pytestmark = pytest.mark.gpu


def small_mix():
    chans = [L.Channel(96, 24, 20, 12, 2, reynolds=80.0), L.Channel(200, 50, 40, 25, 4),
             L.Channel(64, 16, 12, 8, 1, reynolds=300.0), L.Channel(256, 64, 60, 30, 6),
             L.Channel(130, 32, 25, 17, 3, reynolds=40.0), L.Channel(77, 19, 15, 9, 2)]
    steps = [700, 1200, 300, 900, 1500, 250]
    stor = ["f64", "f32", "f64", "f32", "f64", "f32"]
    return [Request(i, 0.002 * i, c, s, st) for i, (c, s, st) in enumerate(zip(chans, steps, stor))]


def test_results_same_alone_and_in_mix():
    reqs = small_mix()
    mixed = SolveServer(max_active=4, slice_ms=1.0, graph_steps=25, keep_fields=True).run(reqs)
    assert [r.rid for r in mixed] == list(range(len(reqs)))
    for req, got in zip(reqs, mixed):
        alone = SolveServer(max_active=1, slice_ms=50.0, graph_steps=7,
                            keep_fields=True).run([Request(0, 0.0, req.channel, req.steps,
                                                           req.storage)])[0]
        for a, b in zip(got.fields, alone.fields):
            assert np.array_equal(a, b), req.rid
        assert np.array_equal(got.diagnostics, alone.diagnostics), req.rid
        # and the same as a plain context stepping straight through
        with L.Lattice(req.channel, storage=req.storage) as lat:
            lat.init_equilibrium()
            lat.step(req.steps, "combine")
            direct = lat.macros()
        for a, b in zip(got.fields, direct):
            assert np.array_equal(a, b), req.rid


def test_admission_budget_and_slots():
    reqs = small_mix()
    fps = [L.footprint_of(r.channel, r.storage) for r in reqs]
    budget = sorted(fps)[-1] + sorted(fps)[-2]   # at most ~two big solves at once
    srv = SolveServer(hbm_budget=budget, max_active=3, slice_ms=1.0, graph_steps=25)
    res = srv.run(reqs)
    assert len(res) == len(reqs)
    assert srv.stats["peak_bytes"] <= budget
    assert srv.stats["peak_active"] <= 3
    assert any(r.queued > 0.0 for r in res)
    s = summarize(res)
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_admission_counts_the_sweeps_buffers():
    """A 2048 x 1024 FP64 solve runs the six-step sweep, whose two further
    buffers the admission counts; the context then holds exactly that."""
    ch = L.Channel.geometry(2048, 1024, "B")
    req = Request(0, 0.0, ch, 12, "f64")
    srv = SolveServer(max_active=1, slice_ms=5.0, graph_steps=20)
    assert srv._variant(req) == "panel6"
    srv.run([req])
    want = L.footprint_of(ch, "f64", "panel6")
    assert srv.stats["peak_bytes"] == want > L.footprint_of(ch, "f64")
    with L.Lattice(ch) as lat:
        lat.init_equilibrium()
        lat.step(12, "panel6")
        assert lat.footprint() == want


def test_late_arrival_overlaps_running_solve():
    big = Request(0, 0.0, L.Channel(2048, 512, 400, 256, 40), 20000, "f64")
    small = Request(1, 0.05, L.Channel(64, 16, 12, 8, 2), 200, "f32")
    res = SolveServer(max_active=4, slice_ms=5.0, graph_steps=20).run([big, small])
    # the small solve is admitted while the big one runs and finishes first
    assert res[1].admitted < res[0].finished
    assert res[1].finished < res[0].finished
