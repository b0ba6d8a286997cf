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
    """A 2048 x 1024 FP64 solve runs the six-step sweep, whose three further
    buffers the admission counts (two for its six-step launches, a third for
    the eight-step launch a remainder of 2 or 3 steps takes); a context holds
    two after 12 steps (1 + 6 + 4 + 1) and all three after 9 (1 + 8)."""
    ch = L.Channel.geometry(2048, 1024, "B")
    req = Request(0, 0.0, ch, 12, "f64")
    srv = SolveServer(max_active=1, slice_ms=5.0, graph_steps=20)
    assert srv._variant(req) == "panel6"
    srv.run([req])
    want = L.footprint_of(ch, "f64", "panel6")
    assert srv.stats["peak_bytes"] == want > L.footprint_of(ch, "f64")
    buf = (want - L.footprint_of(ch, "f64")) // 3
    for steps, held in ((12, want - buf), (9, want)):
        with L.Lattice(ch) as lat:
            lat.init_equilibrium()
            lat.step(steps, "panel6")
            assert lat.footprint() == held, steps


def test_late_arrival_overlaps_running_solve():
    big = Request(0, 0.0, L.Channel(2048, 512, 400, 256, 40), 20000, "f64")
    small = Request(1, 0.05, L.Channel(64, 16, 12, 8, 2), 200, "f32")
    res = SolveServer(max_active=4, slice_ms=5.0, graph_steps=20).run([big, small])
    # the small solve is admitted while the big one runs and finishes first
    assert res[1].admitted < res[0].finished
    assert res[1].finished < res[0].finished


def test_pool_contexts_equal_plain_contexts():
    """Contexts allocated from the stream-ordered pool give the same bits as
    hipMalloc'd ones, a destroyed context's memory is reused, and the pool
    refuses to go while a context still uses it."""
    ch = L.Channel(300, 75, 60, 37, 6, reynolds=120.0)
    want = {}
    for st in ("f64", "f32"):
        with L.Lattice(ch, storage=st) as lat:
            lat.init_equilibrium()
            lat.step(37, "panel")
            want[st] = lat.macros()
    L.pool_reserve(64 << 20)
    try:
        for rep in range(2):
            for st in ("f64", "f32"):
                with L.Lattice(ch, storage=st) as lat:
                    lat.init_equilibrium()
                    lat.step(37, "panel")
                    got = lat.macros()
                    if rep == 1 and st == "f64":
                        with pytest.raises(L.LbmError):
                            L.pool_release()
                for a, b in zip(got, want[st]):
                    assert np.array_equal(a, b), (rep, st)
    finally:
        L.pool_release()
    L.pool_release()    # no pool left: a no-op


def test_macros_in_column_chunks():
    """A lattice larger than the moments scratch (4 M nodes) is read back in
    whole-column chunks (4200 x 1100: 3813 columns, then 387): the fields
    equal computeMacros (lb.c:134-141, the same grouping and true division)
    evaluated in numpy on the downloaded post-stream populations, bit for
    bit, in every column."""
    ch = L.Channel.geometry(4200, 1100, "B")
    with L.Lattice(ch) as lat:
        lat.init_equilibrium()
        lat.step(3, "combine")
        rho, ux, uy = lat.macros()
        f = lat.download()[0][1:-1, 1:-1, :]
    f0, f1, f2, f3, f4, f5, f6, f7, f8 = (f[..., i] for i in range(9))
    north = (f2 + f5) + f6
    mid = (f0 + f1) + f3
    south = (f4 + f7) + f8
    r = (north + mid) + south
    assert np.array_equal(rho, r)
    assert np.array_equal(ux, (((f1 + f5) + f8) - ((f3 + f6) + f7)) / r)
    assert np.array_equal(uy, (north - south) / r)


def _busy_context():
    """A context with about a second of queued work (1800 steps of the
    six-step sweep on 16384 x 4096)."""
    lat = L.Lattice(L.Channel.geometry(16384, 4096, "B"))
    lat.init_equilibrium()
    lat.step(7, "panel6")
    lat.sync()
    lat.step(1800, "panel6")
    return lat


def test_contexts_come_and_go_while_others_run():
    """Admission and completion in the server never wait for the solves
    already running: with the device's context pool, creating a context,
    initialising it, stepping it, reading its fields and destroying it all
    return while another context's queued sweep is still running (that
    context reports busy afterwards).  lbm_hip_create / init / macros /
    diagnostics wait for their own stream only."""
    import time
    L.pool_reserve(256 << 20)
    try:
        big = _busy_context()
        try:
            t0 = time.perf_counter()
            for st in ("f64", "f32"):
                with L.Lattice(L.Channel(300, 75, 60, 37, 6), storage=st) as lat:
                    lat.init_equilibrium()
                    lat.step(20, "combine")
                    lat.macros()
                    lat.diagnostics()
            dt = time.perf_counter() - t0
            assert not big.done(), "the long sweep had finished: the test proves nothing"
            big.sync()
            t_big = time.perf_counter() - t0
            assert dt < 0.5 * t_big, (dt, t_big)
        finally:
            big.close()
    finally:
        L.pool_release()
