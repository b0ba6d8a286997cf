"""Parity on the 16384x4096 bandwidth lattice that bench.py measures: disc
(lx/5, ly/2, ly/10) = (3276, 2048, 409), Re 100, FP64 populations, the
reference's boundary operators and outlet (unsteady.c:84-109, 162-214).

Every sweep the bench times runs 24 steps from the initial equilibrium, in
three calls of 13, 8 and 3 steps: after the first collision the deep sweeps
launch at every depth they have (panel6: two six-step launches, then one
eight-step launch in place of a six- and a two-step one, then a two-step
launch and a single step; panel4: three four-step launches, then two, then a
two-step launch and a single step).  Each result is compared with the
oracle's (the reference's origin order, lb.c:172-341, restated in C) on all
67,108,864 interior nodes, bit for bit: the sweeps this lattice exercises
and the small-lattice tests do not are its 13,000-67,000 interior waves per
launch, the XCD renumbering over that many workgroups, byte offsets above
4 GiB into the population planes, and the deep sweeps' frames along 66-76
bands and 171-260 chunks of 63-96 columns (lb.c:353-437, 936-1069,
1402-1538).  The FP32 and FP16 forms of the deep sweeps, at the chunk widths
they take here, are compared with their one-step sweep the same way."""
import numpy as np
import pytest

import lbm_amd as L
from oracle_lib import Channel as OChannel

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

LX, LY = 16384, 4096
CALLS = (13, 8, 3)


@pytest.fixture(scope="module")
def reference(orc):
    ch = OChannel.geometry(LX, LY, "B")
    ran, lat, tmp, prho = orc.run(ch, sum(CALLS), "origin")
    assert ran == sum(CALLS)
    return ch, lat, tmp, prho


def same(a, b):
    return np.array_equal(a[1:-1, 1:-1, :], b[1:-1, 1:-1, :])


@pytest.mark.parametrize("variant", ["combine", "origin", "tight", "tight_block", "panel",
                                     "panel4", "panel6"])
def test_bench_lattice_matches_oracle(reference, variant):
    ch, lref, tref, pref = reference
    dev = L.Channel(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r)
    with L.Lattice(dev) as lat:
        lat.kernel_timing(True)
        lat.init_equilibrium()
        for n in CALLS:
            lat.step(n, variant)
        recs = lat.kernel_times()
        lo, to, po = lat.download()
    depths = {r["depth"] for r in recs}
    want = {"origin": {1}, "combine": {1}, "tight": {1}, "tight_block": {1, 2},
            "panel": {1, 2}, "panel4": {1, 2, 4}, "panel6": {1, 2, 6, 8}}[variant]
    assert depths == want, recs
    ok_l = same(lo, lref)
    ok_t = same(to, tref)
    if not (ok_l and ok_t):
        d = float(np.abs(lo[1:-1, 1:-1] - lref[1:-1, 1:-1]).max())
        pytest.fail(f"{variant}: differs from the oracle on 16384x4096 (max-abs {d:g}, "
                    f"lattice equal {ok_l}, tmpLattice equal {ok_t})")
    assert np.array_equal(po[2:LY], pref[2:LY])


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_bench_lattice_reduced_deep_sweeps_equal_single_steps(storage):
    """FP32 / FP16 storage on the same lattice and step calls: panel4 and
    panel6 (96-column chunks here, frame strips of 16 columns; panel6's
    remainder as an eight-step launch) give the bits of the one-step sweep (FP32: the same stochastic rounding
    per collision index)."""
    dev = L.Channel.geometry(LX, LY, "B")

    def run(variant):
        with L.Lattice(dev, storage=storage) as lat:
            lat.init_equilibrium()
            for n in CALLS:
                lat.step(n, variant)
            return lat.download()

    ref = run("combine")
    for variant in ("panel4", "panel6"):
        got = run(variant)
        for name, a, b in zip(("lattice", "tmpLattice", "outlet rho"), got, ref):
            assert np.array_equal(a, b), f"{storage} {variant}: {name} differs from combine"
        del got
