"""makePeriodic on the device (lbm_hip_make_periodic, lb.c:1783-1817), through
the C-ABI: against the trajectory the real reference wrote (copy after step 5
of 12), and against the oracle for every sweep before and after the copy.
FP64 is bit-exact; the reduced formats copy stored values exactly."""
import json
import os

import numpy as np
import pytest

import lbm_amd as L
from oracle_lib import Channel as OChannel

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def dev_channel(c: OChannel) -> L.Channel:
    return L.Channel(c.lx, c.ly, c.obst_x, c.obst_y, c.obst_r, c.u_max, c.reynolds,
                     c.bc_family, c.zgb)


def inner(a):
    return a[1:-1, 1:-1, :]


def test_periodic_matches_reference_trajectory():
    g = np.load(os.path.join(GOLD, "periodic.npz"))
    ch = OChannel(**json.loads(str(g["channel"])))
    after, steps = int(g["traj_after"]), int(g["traj_steps"])
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        lat.step(after, "origin")
        lat.make_periodic()
        lat.step(steps - after, "origin")
        lo, to, po = lat.download()
    assert np.array_equal(inner(lo), inner(g["lat"]))
    assert np.array_equal(inner(to), inner(g["tmp"]))
    assert np.array_equal(po[2:ch.ly], g["prho"][2:ch.ly])


@pytest.mark.parametrize("before,after_v,k,m,size", [
    ("origin", "origin", 0, 6, (64, 40)),
    ("origin", "combine", 5, 7, (64, 40)),
    ("combine", "combine", 4, 9, (64, 40)),
    ("combine", "panel", 3, 10, (600, 400)),
    ("panel4", "panel4", 8, 9, (600, 400)),
    ("tight_block", "combine", 6, 3, (600, 400)),
])
def test_periodic_against_oracle(orc, before, after_v, k, m, size):
    """k steps of one sweep, the copy, m steps of another (every sweep gives
    the same bits, so the oracle runs origin throughout)."""
    ch = OChannel.geometry(size[0], size[1], "B")
    _, lref, tref, pref = orc.run_periodic(ch, k + m, k) if k else (None,) * 4
    if not k:
        _, l0, t0, p0 = orc.run(ch, 0, "origin")
        l0 = orc.make_periodic(ch.lx, ch.ly, l0)
        _, lref, tref, pref = orc.run_from(ch, m, l0, t0, p0, "origin")
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        if k:
            lat.step(k, before)
        lat.make_periodic()
        lat.step(m, after_v)
        lo, to, po = lat.download()
    what = f"{before} {k} / copy / {after_v} {m} on {size}"
    assert np.array_equal(inner(lo), inner(lref)), what
    assert np.array_equal(inner(to), inner(tref)), what
    assert np.array_equal(po[2:ch.ly], pref[2:ch.ly]), what


@pytest.mark.parametrize("storage,narrow", [("f32", np.float32), ("f16", np.float16)])
def test_periodic_reduced_storage_copies_stored_values(storage, narrow):
    """In FP32/FP16 storage the copy moves stored deviations unchanged: each
    written slot of the post-stream state equals the post-collision value of
    the interior node it wraps around from, and a slot no interior node feeds
    holds the stored form of 0 (-w_i rounded to the format)."""
    ch = L.Channel(96, 48, 20, 24, 5)
    lx, ly = ch.lx, ch.ly
    with L.Lattice(ch, storage=storage) as lat:
        lat.init_equilibrium()
        lat.step(7, "combine")
        lat.make_periodic()
        lo, to, _ = lat.download()
    zero = np.array([float(narrow(-w)) + w for w in W])

    def ghost(i, gx, gy):
        sx, sy = gx - CX[i], gy - CY[i]
        return to[sx, sy, i] if 1 <= sx <= lx and 1 <= sy <= ly else zero[i]

    for x in range(1, lx + 1):
        if x not in (1, lx):
            assert lo[x, ly, 7] == ghost(7, x, 0) and lo[x, 1, 6] == ghost(6, x, ly + 1)
            assert lo[x, ly, 8] == ghost(8, x, 0) and lo[x, 1, 5] == ghost(5, x, ly + 1)
        assert lo[x, ly, 4] == ghost(4, x, 0) and lo[x, 1, 2] == ghost(2, x, ly + 1)
    for y in range(2, ly):
        for i in (1, 5, 8):
            assert lo[1, y, i] == ghost(i, lx + 1, y)
        for i in (3, 6, 7):
            assert lo[lx, y, i] == ghost(i, 0, y)
    assert lo[1, 1, 5] == ghost(5, lx + 1, ly + 1) and lo[lx, 1, 6] == ghost(6, 0, ly + 1)
    assert lo[lx, ly, 7] == ghost(7, 0, 0) and lo[1, ly, 8] == ghost(8, lx + 1, 0)

