"""Batches: every channel of a persistent batched launch evolves bit for bit
as the same channel in its own context (and as the oracle), for every launch
shape, with and without hipGraph replay."""
import numpy as np
import pytest

import lbm_amd as L
from lbm_amd.batch import Batch, autotune, channels_from_seed

pytestmark = pytest.mark.gpu


def small_channels():
    # uneven sizes, a channel shorter than one tile row, discs of radius 1..6
    return [L.Channel(130, 33, 30, 16, 3, reynolds=150.0),
            L.Channel(17, 5, 4, 2, 1, reynolds=40.0),
            L.Channel(64, 16, 12, 8, 2),
            L.Channel(300, 75, 60, 37, 6, reynolds=20.0),
            L.Channel(48, 300, 10, 150, 5, reynolds=60.0),
            L.Channel(9, 4, 2, 2, 1)]


def alone(ch, steps, storage):
    with L.Lattice(ch, storage=storage) as lat:
        lat.init_equilibrium()
        lat.step(steps, "combine")
        return lat.download()


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_batch_equals_single_contexts(storage):
    """Bit-identical for every storage format: the batched and the
    single-lattice kernels run the same node step."""
    chans = small_channels()
    steps = 23
    with Batch(chans, storage) as b:
        b.init_equilibrium()
        b.step(steps)
        got = [b.download(i) for i in range(len(chans))]
    for i, ch in enumerate(chans):
        want = alone(ch, steps, storage)
        for g, w, name in zip(got[i], want, ("lattice", "tmp", "outlet")):
            a, b_ = (g[2:ch.ly], w[2:ch.ly]) if name == "outlet" else (g[1:-1, 1:-1], w[1:-1, 1:-1])
            assert np.array_equal(a, b_), (i, name, np.abs(a - b_).max())


def test_batch_matches_oracle(orc):
    from oracle_lib import Channel as OC
    chans = small_channels()[:4]
    with Batch(chans, "f64") as b:
        b.init_equilibrium()
        b.step(40)
        got = [b.download(i)[0] for i in range(len(chans))]
    for ch, lat in zip(chans, got):
        oc = OC(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r, ch.u_max, ch.reynolds)
        _, lref, _, _ = orc.run(oc, 40)
        assert np.array_equal(lat[1:-1, 1:-1], lref[1:-1, 1:-1])


@pytest.mark.parametrize("ty_cap,wgs", [(32, 1), (64, 3), (128, 16), (256, 2)])
def test_launch_shape_does_not_change_results(ty_cap, wgs):
    chans = small_channels()
    with Batch(chans, "f64") as a, Batch(chans, "f64") as b:
        b.set_launch(wgs, ty_cap, 0)
        for x in (a, b):
            x.init_equilibrium()
            x.step(17)
        for i in range(len(chans)):
            assert np.array_equal(a.download(i)[0], b.download(i)[0])


@pytest.mark.parametrize("graph_steps", [1, 4, 7])
def test_graph_replay_equals_eager(graph_steps):
    """Several replays of the captured step sequence, with eager steps and
    host synchronisation between them, match a purely eager run."""
    chans = small_channels()
    with Batch(chans, "f64") as eager, Batch(chans, "f64") as graph:
        graph.set_launch(0, 0, graph_steps)
        for x in (eager, graph):
            x.init_equilibrium()
        for chunk in (1, 2 * graph_steps, 3, graph_steps, 5 * graph_steps + 1):
            eager.step(chunk)
            graph.step(chunk)
            graph.sync()
            for i in (0, 3, 5):
                assert np.array_equal(eager.download(i)[0], graph.download(i)[0]), (chunk, i)
        assert eager.steps_done() == graph.steps_done()


def test_seeded_workload_and_autotune(tmp_path):
    chans = channels_from_seed(32, seed=3)
    assert len(chans) == 32
    for c in chans:
        assert 128 <= c.lx <= 2048 and c.ly == c.lx // 4
        assert 0.15 * c.lx - 1 <= c.obst_x <= 0.3 * c.lx
        assert 20.0 <= c.reynolds <= 200.0 and c.obst_r >= 1
    with Batch(chans, "f32") as b:
        b.init_equilibrium()
        res = autotune(b, steps=4, graph_steps=2, ty_caps=(64, 256), wgs=(4, 8),
                       cache_dir=str(tmp_path))
        assert res["glups"] > 0 and len(res["candidates"]) == 4
        again = autotune(b, cache_dir=str(tmp_path))
        assert again.get("cached") and again["ty_cap"] == res["ty_cap"]
