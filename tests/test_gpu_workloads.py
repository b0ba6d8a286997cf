"""The batching and serving workloads at their own sizes.

Batching: the 512 channels of seed 0 (lx log-uniform in [128, 2048],
ly = lx/4, 96.9 M cells, 118 of them with tau < 0.505), advanced with
10-step hipGraphs and every multi-step sweep (two steps: LDS tiles, register
march; four and six: the deep register marches with two-step remainders),
as the bench runs them.  The
smallest and the largest channel and those with tau < 0.505 equal the same
channel alone in its own context bit for bit, in every storage format, and
three channels (smallest, median, largest) equal the oracle (the reference's
origin order, lb.c:172-341, unsteady.c:424-531).

Serving: the seeded Poisson stream of the serving workload (40 solves, rate
2/s, seed 0: lx log-uniform in [256, 4096], Re in [20, 500], FP64 / FP32)
with its step counts and arrival times cut by 50 and 10 so the test runs in
seconds; every solve's fields and diagnostics equal the same solve run alone
on its own context with the same sweep, bit for bit."""
import dataclasses

import numpy as np
import pytest

import lbm_amd as L
from lbm_amd.batch import Batch, channels_from_seed
from lbm_amd.serve import SolveServer, poisson_requests

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

STEPS = 21      # one single step after initialisation, then two 10-step graphs


@pytest.fixture(scope="module")
def seeded():
    chans = channels_from_seed(512, seed=0)
    cells = [c.lx * c.ly for c in chans]
    tau = [i for i, c in enumerate(chans) if 1.0 / c.omega() < 0.505]
    order = np.argsort(cells, kind="stable")
    return chans, int(order[0]), int(order[len(order) // 2]), int(order[-1]), tau


def alone(ch, steps, storage):
    with L.Lattice(ch, storage=storage) as lat:
        lat.init_equilibrium()
        lat.step(steps, "combine")
        return lat.download()


def test_seeded_batch_shape(seeded):
    chans, small, _, large, tau = seeded
    assert len(chans) == 512
    assert sum(c.lx * c.ly for c in chans) == 96856364
    assert len(tau) == 118
    assert chans[small].lx < 140 and chans[large].lx > 1900


@pytest.mark.parametrize("sweep", ["tight_block", "panel", "panel4", "panel6"])
@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_seeded_batch_equals_single_contexts(seeded, storage, sweep):
    chans, small, _, large, tau = seeded
    ids = [small, large] + (tau if storage == "f64" else tau[:12])
    with Batch(chans, storage) as b:
        b.set_sweep(sweep)
        b.set_launch(0, 0, 10)
        b.init_equilibrium()
        b.step(STEPS)
        assert b.steps_done() == STEPS
        got = {i: b.download(i) for i in ids}
    for i in ids:
        ch = chans[i]
        want = alone(ch, STEPS, storage)
        for g, w, name in zip(got[i], want, ("lattice", "tmp", "outlet")):
            a, b_ = (g[2:ch.ly], w[2:ch.ly]) if name == "outlet" else (g[1:-1, 1:-1], w[1:-1, 1:-1])
            assert np.array_equal(a, b_), (storage, i, ch, name)


@pytest.mark.parametrize("sweep", ["tight_block", "panel", "panel4", "panel6"])
def test_seeded_batch_matches_oracle(seeded, orc, sweep):
    from oracle_lib import Channel as OC
    chans, small, median, large, _ = seeded
    ids = [small, median, large]
    with Batch(chans, "f64") as b:
        b.set_sweep(sweep)
        b.set_launch(0, 0, 10)
        b.init_equilibrium()
        b.step(STEPS)
        got = {i: b.download(i) for i in ids}
    for i in ids:
        ch = chans[i]
        oc = OC(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r, ch.u_max, ch.reynolds)
        _, lref, tref, pref = orc.run(oc, STEPS)
        lat, tmp, prho = got[i]
        assert np.array_equal(lat[1:-1, 1:-1], lref[1:-1, 1:-1]), (i, ch)
        assert np.array_equal(tmp[1:-1, 1:-1], tref[1:-1, 1:-1]), (i, ch)
        assert np.array_equal(prho[2:ch.ly], pref[2:ch.ly]), (i, ch)


def test_seeded_poisson_stream_each_solve_as_alone():
    reqs = [dataclasses.replace(r, steps=max(20, r.steps // 50), arrival=r.arrival / 10)
            for r in poisson_requests(40, 2.0, seed=0)]
    srv = SolveServer(max_active=8, slice_ms=5.0, graph_steps=50, keep_fields=True)
    res = srv.run(reqs)
    assert [r.rid for r in res] == list(range(40))
    assert srv.stats["peak_active"] >= 2          # solves overlapped on the device
    assert {r.storage for r in res} == {"f64", "f32"}
    for req, got in zip(reqs, res):
        with L.Lattice(req.channel, storage=req.storage) as lat:
            lat.init_equilibrium()
            lat.step(req.steps, srv._variant(req))
            fields = lat.macros()
            diag = lat.diagnostics()
        for a, b in zip(got.fields, fields):
            assert np.array_equal(a, b), (req.rid, req.channel, req.storage)
        assert np.array_equal(got.diagnostics, diag), req.rid
