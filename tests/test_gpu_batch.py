"""Batches: every channel of a batched launch evolves bit for bit as the same
channel in its own context (and as the oracle), for one, two, four and six
steps per launch, every launch shape, tile and chunk width, with and without
hipGraph replay."""
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


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
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
            x.set_depth(1)
            x.init_equilibrium()
            x.step(17)
        for i in range(len(chans)):
            assert np.array_equal(a.download(i)[0], b.download(i)[0])


@pytest.mark.parametrize("storage,tile_cols", [("f64", 3), ("f64", 8), ("f64", 33),
                                               ("f32", 3), ("f32", 14), ("f32", 69),
                                               ("f16", 10), ("f16", 40)])
def test_two_steps_per_launch_equal_one(storage, tile_cols):
    """The two-step LDS tiles (the default) against the one-step launch,
    over odd and even step counts and several tile widths (3 columns: every
    channel's right end in a tile of its own width; 33 / 69: the widest tile
    LDS holds with FP64 / FP32 values)."""
    chans = small_channels()
    with Batch(chans, storage) as one, Batch(chans, storage) as two:
        one.set_depth(1)
        two.set_depth(2, tile_cols)
        for x in (one, two):
            x.init_equilibrium()
        for chunk in (1, 2, 5, 12):
            one.step(chunk)
            two.step(chunk)
            for i in range(len(chans)):
                g, w = two.download(i), one.download(i)
                ch = chans[i]
                assert np.array_equal(g[0][1:-1, 1:-1], w[0][1:-1, 1:-1]), (chunk, i)
                assert np.array_equal(g[1][1:-1, 1:-1], w[1][1:-1, 1:-1]), (chunk, i)
                assert np.array_equal(g[2][2:ch.ly], w[2][2:ch.ly]), (chunk, i)
        assert one.steps_done() == two.steps_done() == 20


def test_two_step_tile_width_limits():
    with Batch(small_channels()[:2], "f64") as b, Batch(small_channels()[:2], "f16") as h:
        with pytest.raises(L.LbmError, match="tile_cols"):
            b.set_depth(2, 34)
        with pytest.raises(L.LbmError, match="tile_cols"):
            h.set_depth(2, 70)
        with pytest.raises(L.LbmError, match="tile_cols"):
            h.set_depth(2, 2)
        b.set_depth(2, 33)
        h.set_depth(2, 69)


def test_narrow_channel_keeps_one_step_per_launch():
    """A channel with lx = 3 has no room for the two-step tile's outlet
    extrapolation (as the single-lattice two-step sweeps): such a batch
    defaults to one step per launch, refuses two, and stays bit-exact."""
    chans = small_channels()[:3] + [L.Channel(3, 5, 1, 2, 1)]
    with Batch(chans, "f64") as b:
        assert b.launch["depth"] == 1
        with pytest.raises(L.LbmError, match="lx >= 4"):
            b.set_depth(2)
        b.init_equilibrium()
        b.step(9)
        got = [b.download(i)[0] for i in range(len(chans))]
    for i, ch in enumerate(chans):
        assert np.array_equal(got[i][1:-1, 1:-1], alone(ch, 9, "f64")[0][1:-1, 1:-1]), i


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
                       tile_cols=(18,), panel_cols=(6, 8), deep_cols=(8,),
                       cache_dir=str(tmp_path))
        assert res["glups"] > 0 and len(res["candidates"]) == 9
        assert [c["depth"] for c in res["candidates"]] == [1, 1, 1, 1, 2, 2, 2, 2, 2]
        assert [c["sweep"] for c in res["candidates"]][-5:] == [
            "tight_block", "panel", "panel", "panel4", "panel6"]
        assert [c["chunk_cols"] for c in res["candidates"]][-4:] == [6, 8, 8, 8]
        assert [c["graph_steps"] for c in res["candidates"]][-3:] == [2, 12, 12]
        assert res["steps_per_launch"] == {"panel4": 4, "panel6": 6}.get(
            res["sweep"], res["depth"])
        again = autotune(b, cache_dir=str(tmp_path))
        assert again.get("cached") and again["ty_cap"] == res["ty_cap"]
        assert again["depth"] == res["depth"] and b.launch["depth"] == res["depth"]
        assert again["sweep"] == res["sweep"] and b.launch["sweep"] == res["sweep"]
        assert again["chunk_cols"] == res["chunk_cols"]


def test_autotune_fp64_times_both_depths(tmp_path):
    """The multi-step candidates of FP64 widths; the state keeps advancing."""
    chans = channels_from_seed(16, seed=5)
    with Batch(chans, "f64") as b:
        b.init_equilibrium()
        res = autotune(b, steps=4, graph_steps=3, ty_caps=(64,), wgs=(4,), tile_cols=(8, 14),
                       panel_cols=(6,), deep_cols=(12,), deep_graph_steps=6,
                       cache_dir=str(tmp_path))
        assert [c["depth"] for c in res["candidates"]] == [1, 2, 2, 2, 2, 2]
        assert [c["sweep"] for c in res["candidates"]][-3:] == ["panel", "panel4", "panel6"]
        assert res["graph_steps"] == (6 if res["sweep"] in ("panel4", "panel6") else 3)
        assert res["depth"] in (1, 2) and b.launch["depth"] == res["depth"]
        b.step(5)
        assert b.steps_done() > 0


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
@pytest.mark.parametrize("graph_steps", [0, 6])
def test_batch_register_march_equals_single_contexts(storage, graph_steps):
    """The two-step register march over every channel's wave grid (interior
    waves from one flat work table, the frame around them in two one-step
    passes through the third pool) gives every channel the bits of the
    channel alone, eagerly and replayed from hipGraphs."""
    chans = small_channels() + [L.Channel(700, 250, 140, 125, 25, reynolds=90.0)]
    steps = 31
    with Batch(chans, storage) as b:
        b.set_sweep("panel")
        b.set_launch(0, 0, graph_steps)
        b.init_equilibrium()
        b.step(steps)
        got = [b.download(i) for i in range(len(chans))]
    for i, ch in enumerate(chans):
        want = alone(ch, steps, storage)
        for g, w, name in zip(got[i], want, ("lattice", "tmp", "outlet")):
            a, b_ = (g[2:ch.ly], w[2:ch.ly]) if name == "outlet" else (g[1:-1, 1:-1], w[1:-1, 1:-1])
            assert np.array_equal(a, b_), (i, name, storage, np.abs(a - b_).max())


@pytest.mark.parametrize("chunk_cols", [4, 16, 64])
def test_batch_register_march_chunk_widths(chunk_cols):
    """Every chunk width cuts the channels into other interior boxes and
    frames (the widest leaves the small channels all frame): same bits."""
    chans = small_channels() + [L.Channel(700, 250, 140, 125, 25, reynolds=90.0),
                                L.Channel(1030, 130, 200, 64, 20)]
    steps = 17
    with Batch(chans, "f32") as b:
        b.set_sweep("panel", chunk_cols)
        b.init_equilibrium()
        b.step(steps)
        got = [b.download(i) for i in range(len(chans))]
    for i, ch in enumerate(chans):
        want = alone(ch, steps, "f32")
        for g, w, name in zip(got[i], want, ("lattice", "tmp", "outlet")):
            a, b_ = (g[2:ch.ly], w[2:ch.ly]) if name == "outlet" else (g[1:-1, 1:-1], w[1:-1, 1:-1])
            assert np.array_equal(a, b_), (i, name, chunk_cols)


# channels with a deep interior for the four- and six-step marches (at least
# 124 interior rows once the halo rows are off the walls)
DEEP = [L.Channel(700, 250, 140, 125, 25, reynolds=90.0),
        L.Channel(1030, 140, 200, 70, 20),
        L.Channel(520, 300, 100, 150, 30, reynolds=300.0)]


def assert_as_alone(got, chans, steps, storage, tag):
    for i, ch in enumerate(chans):
        want = alone(ch, steps, storage)
        for g, w, name in zip(got[i], want, ("lattice", "tmp", "outlet")):
            a, b_ = (g[2:ch.ly], w[2:ch.ly]) if name == "outlet" else (g[1:-1, 1:-1], w[1:-1, 1:-1])
            assert np.array_equal(a, b_), (i, name, tag, np.abs(a - b_).max())


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
@pytest.mark.parametrize("sweep", ["panel4", "panel6"])
@pytest.mark.parametrize("graph_steps", [0, 12])
def test_batch_deep_march_equals_single_contexts(storage, sweep, graph_steps):
    """The four- and six-step register marches over the channels with a deep
    interior (interior waves in one launch, the frame in two-step passes and
    the rings inside it through further buffers), two-step passes for the
    channels without one, two-step launches for the remainders (35 steps:
    one, then four-step launches and a two-step one, or six-step launches
    and two two-step ones): every channel keeps the bits of the channel
    alone, eagerly and replayed from hipGraphs."""
    chans = small_channels() + DEEP
    steps = 35
    with Batch(chans, storage) as b:
        base = b.footprint()
        b.set_sweep(sweep)
        # one (four steps) or two (six) further buffers of a pool's size
        nmid = 1 if sweep == "panel4" else 2
        assert abs((b.footprint() - base) - nmid * base / 2) < 0.01 * base
        b.set_launch(0, 0, graph_steps)
        b.init_equilibrium()
        b.step(steps)
        assert b.steps_done() == steps
        got = [b.download(i) for i in range(len(chans))]
    assert_as_alone(got, chans, steps, storage, (sweep, graph_steps))


@pytest.mark.parametrize("chunk_cols", [4, 24, 64])
@pytest.mark.parametrize("sweep", ["panel4", "panel6"])
def test_batch_deep_march_chunk_widths(sweep, chunk_cols):
    """Every deep chunk width cuts the deep channels into other interior
    boxes, frames and rings: same bits."""
    chans = small_channels()[:3] + DEEP
    steps = 25
    with Batch(chans, "f32") as b:
        b.set_sweep(sweep, chunk_cols)
        b.init_equilibrium()
        b.step(steps)
        got = [b.download(i) for i in range(len(chans))]
    assert_as_alone(got, chans, steps, "f32", (sweep, chunk_cols))


def test_batch_deep_sweep_settings():
    """Odd or out-of-range chunk widths are refused and leave the sweep as it
    was; a batch holding a channel narrower than four columns refuses the
    register marches; a batch whose channels are all too small for the deep
    march still runs LBM_PANEL4 (as two-step passes) with the same bits."""
    chans = small_channels()[:3] + DEEP[:1]
    with Batch(chans, "f64") as b:
        b.set_sweep("panel6", 32)
        for bad in (5, 2, 66):
            with pytest.raises(L.LbmError, match="chunk_cols"):
                b.set_sweep("panel6", bad)
        b.init_equilibrium()
        b.step(13)
        got = [b.download(i) for i in range(len(chans))]
    assert_as_alone(got, chans, 13, "f64", "after refused settings")
    with Batch([L.Channel(3, 8, 1, 4, 1)] + DEEP[:1], "f64") as b:
        with pytest.raises(L.LbmError):
            b.set_sweep("panel4")
    # fewer than 12 rows: below the four-step frame's two widened strips
    small = [L.Channel(17, 5, 4, 2, 1, reynolds=40.0), L.Channel(9, 4, 2, 2, 1),
             L.Channel(40, 10, 8, 5, 2)]
    with Batch(small, "f32") as b:
        b.set_sweep("panel4", 16)
        b.init_equilibrium()
        b.step(11)
        got = [b.download(i) for i in range(len(small))]
    assert_as_alone(got, small, 11, "f32", "all shallow")


def test_batch_deep_march_matches_oracle(orc):
    """Switching between every multi-step sweep of a batch, the deep marches
    with remainders among them: every channel equals the oracle."""
    from oracle_lib import Channel as OC
    chans = small_channels()[:2] + DEEP[:2]
    with Batch(chans, "f64") as b:
        b.init_equilibrium()
        for n, sweep in ((5, "panel4"), (7, "panel6"), (6, "tight_block"), (9, "panel4"),
                         (4, "panel"), (13, "panel6")):
            b.set_sweep(sweep)
            b.step(n)
        got = [b.download(i)[0] for i in range(len(chans))]
    for ch, lat in zip(chans, got):
        oc = OC(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r, ch.u_max, ch.reynolds)
        _, lref, _, _ = orc.run(oc, 44)
        assert np.array_equal(lat[1:-1, 1:-1], lref[1:-1, 1:-1])


def test_batch_sweep_switches_keep_results(orc):
    """Switching a batch between the LDS tiles and the register march between
    steps changes nothing: every channel still equals the oracle."""
    from oracle_lib import Channel as OC
    chans = small_channels()[:4]
    with Batch(chans, "f64") as b:
        b.init_equilibrium()
        for n, sweep in ((5, "panel"), (6, "tight_block"), (9, "panel"), (4, "tight_block")):
            b.set_sweep(sweep)
            b.step(n)
        got = [b.download(i)[0] for i in range(len(chans))]
    for ch, lat in zip(chans, got):
        oc = OC(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r, ch.u_max, ch.reynolds)
        _, lref, _, _ = orc.run(oc, 24)
        assert np.array_equal(lat[1:-1, 1:-1], lref[1:-1, 1:-1])
