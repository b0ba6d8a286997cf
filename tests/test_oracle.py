"""Pinning the CPU oracle: against the golden vectors the real reference wrote
(tests/golden), against the reference harness itself where it was built, and
against what the mathematics fixes (moments of the equilibrium, invariants of
the operators, closed forms of the setup)."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle_lib import BC_NAMES, Channel, reference

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


# ---------------------------------------------------------------- golden KATs

def test_kat_macros_equilibrium_bgk(orc):
    k = np.load(os.path.join(GOLD, "kat.npz"))
    got = np.array([orc.macros(f) for f in k["macros_in"]])
    assert np.array_equal(got, k["macros_out"])
    got = np.array([orc.equilibrium(int(r[0]), r[1], r[2], r[3], r[4]) for r in k["eq_in"]])
    assert np.array_equal(got, k["eq_out"])
    got = np.array([orc.bgk(f, w) for f, w in zip(k["bgk_in"], k["bgk_omega"])])
    assert np.array_equal(got, k["bgk_out"])


@pytest.mark.parametrize("which", range(17))
def test_kat_node_operators(orc, which):
    k = np.load(os.path.join(GOLD, "kat.npz"))
    assert str(k["bc_names"][which]) == BC_NAMES[which]
    for j in range(k["bc_in"].shape[1]):
        got = orc.apply_bc(which, k["bc_in"][which, j], k["bc_p"][which, j, 0],
                           k["bc_p"][which, j, 1], k["bc_omega"][which, j])
        assert np.array_equal(got, k["bc_out"][which, j]), (BC_NAMES[which], j)


@pytest.mark.parametrize("name", ["traj_32x16_reg.npz", "traj_32x16_zh.npz",
                                  "traj_64x32_reg.npz"])
def test_trajectories_bitwise(orc, name):
    g = np.load(os.path.join(GOLD, name))
    ch = Channel(**json.loads(str(g["channel"])))
    steps = sorted(int(k.split("_")[1]) for k in g.files if k.startswith("lat_"))
    for s in steps:
        ran, lat, tmp, prho = orc.run(ch, s, "origin")
        assert ran == s
        assert np.array_equal(lat, g[f"lat_{s}"]), (name, s)   # ghost ring included
        assert np.array_equal(tmp, g[f"tmp_{s}"]), (name, s)
        assert np.array_equal(prho, g[f"prho_{s}"]), (name, s)


def test_make_periodic_against_reference_fixture(orc):
    """makePeriodic (lb.c:1783-1817): on a seeded random lattice, and inside
    a 12-step origin trajectory with the copy after step 5, bit for bit
    against what the real reference wrote (ghost rings included)."""
    g = np.load(os.path.join(GOLD, "periodic.npz"))
    lx, ly = int(g["kat_lx"]), int(g["kat_ly"])
    out = orc.make_periodic(lx, ly, g["kat_in"])
    assert np.array_equal(out, g["kat_out"])
    # every slot the copy writes lies on the perimeter rows / columns
    changed = np.argwhere(out != g["kat_in"])
    assert all(x in (1, lx) or y in (1, ly) for x, y, _ in changed)
    ch = Channel(**json.loads(str(g["channel"])))
    ran, lat, tmp, prho = orc.run_periodic(ch, int(g["traj_steps"]), int(g["traj_after"]))
    assert ran == int(g["traj_steps"])
    assert np.array_equal(lat, g["lat"]) and np.array_equal(tmp, g["tmp"])
    assert np.array_equal(prho, g["prho"])
    # the copy changes the flow (a check that the fixture exercises it)
    assert not np.array_equal(orc.run(ch, ran, "origin")[1], g["lat"])


def test_make_periodic_wraps_each_leaving_population(orc):
    """Closed form of the copy: a population that left through one side
    (and sits in that side's ghost ring) is what the opposite perimeter node
    receives, for every side and corner; the ghost ring itself is untouched."""
    lx, ly = 6, 5
    lat = np.random.default_rng(9).standard_normal((lx + 2, ly + 2, 9))
    out = orc.make_periodic(lx, ly, lat)
    for x in range(1, lx + 1):
        for i in (4, 7, 8):
            if (x, i) not in ((lx, 7), (1, 8)):
                assert out[x, ly, i] == lat[x, 0, i]
        for i in (2, 5, 6):
            if (x, i) not in ((1, 5), (lx, 6)):
                assert out[x, 1, i] == lat[x, ly + 1, i]
    for y in range(1, ly + 1):
        for i in (1, 5, 8):
            if (y, i) not in ((1, 5), (ly, 8)):
                assert out[1, y, i] == lat[lx + 1, y, i]
        for i in (3, 6, 7):
            if (y, i) not in ((1, 6), (ly, 7)):
                assert out[lx, y, i] == lat[0, y, i]
    assert out[1, 1, 5] == lat[lx + 1, ly + 1, 5] and out[lx, 1, 6] == lat[0, ly + 1, 6]
    assert out[lx, ly, 7] == lat[0, 0, 7] and out[1, ly, 8] == lat[lx + 1, 0, 8]
    ring = np.ones((lx + 2, ly + 2), bool)
    ring[1:-1, 1:-1] = False
    assert np.array_equal(out[ring], lat[ring])


@pytest.mark.slow
@pytest.mark.parametrize("geom", ["A", "B"])
def test_c1_hashes(orc, geom):
    rec = json.load(open(os.path.join(GOLD, f"c1_{geom}.json")))
    ch = Channel(**rec["channel"])
    assert ch.omega() == rec["omega"]
    ran, lat, tmp, prho = orc.run(ch, rec["steps"], "origin")
    assert sha(lat) == rec["sha_lattice"]
    assert sha(tmp) == rec["sha_tmp"]
    assert sha(prho) == rec["sha_prho"]


@pytest.mark.parametrize("variant", ["combine", "tight", "tight_block", "panel"])
def test_variant_orders_equal_origin(orc, variant):
    """The five sweep orders of the reference are bit-identical in serial
    (recorded from the reference in golden/variants.json); the oracle's
    restatements of the orders must agree."""
    rec = json.load(open(os.path.join(GOLD, "variants.json")))
    assert all(rec["serial"].values())
    ch = Channel.geometry(64, 32, "A")
    _, l0, t0, p0 = orc.run(ch, 40, "origin")
    _, l1, t1, p1 = orc.run(ch, 40, variant, blk=8)
    assert np.array_equal(l0, l1) and np.array_equal(t0, t1) and np.array_equal(p0, p1)


def test_against_reference_harness(orc):
    ref = reference()
    if not ref.available:
        pytest.skip("reference harness not built here (no reference checkout)")
    for fam in (0, 1):
        ch = Channel.geometry(40, 20, "A", bc_family=fam)
        for v in ("origin", "tight_block", "panel"):
            a = ref.run(ch, 23, v, blk=4)
            b = orc.run(ch, 23, v, blk=4)
            assert a[0] == b[0]
            for x, y in zip(a[1:], b[1:]):
                assert np.array_equal(x, y), (fam, v)


# ------------------------------------------------------------ closed forms

def random_pops(rng, n):
    return W[None, :] * (1.0 + 0.2 * rng.standard_normal((n, 9)))


def test_equilibrium_moments(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        rho, ux, uy = 1 + 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal(), \
            0.1 * rng.standard_normal()
        usq = ux * ux + uy * uy
        f = np.array([orc.equilibrium(i, rho, ux, uy, usq) for i in range(9)])
        assert abs(f.sum() - rho) < 1e-14
        assert abs((CX * f).sum() - rho * ux) < 1e-14
        assert abs((CY * f).sum() - rho * uy) < 1e-14
        assert abs((CX * CX * f).sum() - rho * (1 / 3 + ux * ux)) < 1e-14
        assert abs((CY * CY * f).sum() - rho * (1 / 3 + uy * uy)) < 1e-14
        assert abs((CX * CY * f).sum() - rho * ux * uy) < 1e-14


def test_macros_definition(orc):
    rng = np.random.default_rng(2)
    for f in random_pops(rng, 30):
        r, u, v = orc.macros(f)
        assert abs(r - f.sum()) < 1e-15
        assert abs(u - (CX * f).sum() / f.sum()) < 1e-15
        assert abs(v - (CY * f).sum() / f.sum()) < 1e-15


def test_bgk_conserves_and_fixes_equilibrium(orc):
    rng = np.random.default_rng(3)
    for f in random_pops(rng, 30):
        g = orc.bgk(f, 1.7)
        assert abs(g.sum() - f.sum()) < 1e-14
        assert abs((CX * g).sum() - (CX * f).sum()) < 1e-14
        assert abs((CY * g).sum() - (CY * f).sum()) < 1e-14
        # omega = 1 lands exactly on the equilibrium of the input moments
        r, u, v = orc.macros(f)
        e = np.array([orc.equilibrium(i, r, u, v, u * u + v * v) for i in range(9)])
        assert np.abs(orc.bgk(f, 1.0) - e).max() < 1e-15
        assert np.abs(orc.bgk(e, 1.3) - e).max() < 1e-15


def test_bounce_back_is_involution(orc):
    rng = np.random.default_rng(4)
    for f in random_pops(rng, 10):
        g = orc.apply_bc(0, f, 0, 0, 1.0)
        assert np.array_equal(g, f[OPP])
        assert np.array_equal(orc.apply_bc(0, g, 0, 0, 1.0), f)


SIDES = ["upper", "lower", "left", "right"]
NORMAL = {"upper": (0, 1), "lower": (0, -1), "left": (-1, 0), "right": (1, 0)}


@pytest.mark.parametrize("family", ["ZouHe", "Regularized"])
@pytest.mark.parametrize("side", SIDES)
def test_velocity_operators_impose_velocity(orc, family, side):
    """Before the bulk relaxation, a velocity operator's populations carry
    exactly the prescribed velocity; omega -> 0 exposes that state."""
    which = BC_NAMES.index(f"{side}{family}")
    rng = np.random.default_rng(5)
    for f in random_pops(rng, 10):
        ux, uy = 0.03 * rng.standard_normal(2)
        g = orc.apply_bc(which, f, ux, uy, 1e-300)  # bgk with omega ~ 0 is identity
        r = g.sum()
        assert abs((CX * g).sum() / r - ux) < 1e-13
        assert abs((CY * g).sum() / r - uy) < 1e-13


@pytest.mark.parametrize("family", ["ZouHe", "Regularized"])
@pytest.mark.parametrize("side", SIDES)
def test_pressure_operators_impose_density(orc, family, side):
    which = BC_NAMES.index(f"{side}Pressure{family}")
    rng = np.random.default_rng(6)
    for f in random_pops(rng, 10):
        rho, upar = 1 + 0.02 * rng.standard_normal(), 0.01 * rng.standard_normal()
        g = orc.apply_bc(which, f, rho, upar, 1e-300)
        assert abs(g.sum() - rho) < 1e-13
        nx, ny = NORMAL[side]
        tang = (CX * g).sum() if nx == 0 else (CY * g).sum()
        assert abs(tang / g.sum() - upar) < 1e-13


def test_poiseuille_closed_form():
    ch = Channel.geometry(256, 128, "B")
    from oracle_lib import oracle
    lib = oracle().lib
    lib.orc_poiseuille.restype = __import__("ctypes").c_double
    lib.orc_poiseuille.argtypes = [__import__("ctypes").c_int] * 2 + [__import__("ctypes").c_double]
    vals = np.array([lib.orc_poiseuille(y, ch.ly, ch.u_max) for y in range(1, ch.ly + 1)])
    assert vals[0] == 0.0 and vals[-1] == 0.0
    Lw = ch.ly - 1
    # ly-1 = 127 is odd: the maximum sits half a node off the grid
    assert abs(vals.max() - ch.u_max * (1 - 1 / Lw ** 2)) < 1e-16
    assert np.allclose(vals, vals[::-1], rtol=0, atol=1e-18)


def test_quiescent_channel_stays_at_rest(orc):
    """u_max = 0 and no disc: f stays at the weights except at the four
    corners, which read the zeroed second buffer on the first step (the
    stale-slot rule)."""
    ch = Channel(32, 16, 0, 0, -1, u_max=0.0)
    _, lat, tmp, _ = orc.run(ch, 1)
    inner = lat[1:-1, 1:-1]
    away = inner[2:-2, 2:-2]
    assert np.abs(away - W).max() < 1e-15
    corners = [inner[0, 0], inner[-1, 0], inner[0, -1], inner[-1, -1]]
    assert all(np.abs(c - W).max() > 1e-3 for c in corners)


def test_zgb_extrapolation_exact_for_quadratic(orc):
    """The outlet density after one step is the one-sided second-order
    extrapolation (4 rho(lx-1) - rho(lx-2)) / 3 of the post-stream field."""
    ch = Channel.geometry(32, 16, "A")
    _, lat, _, prho = orc.run(ch, 3)
    _, l2, _, _ = orc.run(ch, 2)
    for y in range(2, ch.ly):
        r1 = orc.macros(l2[ch.lx - 1, y])[0]
        r2 = orc.macros(l2[ch.lx - 2, y])[0]
        assert prho[y] == 4.0 / 3.0 * r1 - 1.0 / 3.0 * r2


def test_omega_values():
    assert Channel.geometry(256, 128, "B").omega() == pytest.approx(1.94401244167963, abs=1e-13)
    assert Channel.geometry(256, 128, "A").omega() == pytest.approx(1.93948797517455, abs=1e-13)
    assert Channel.geometry(16384, 4096, "B").omega() == pytest.approx(1.00928542591845, abs=1e-13)


# --------------------------------------------------------------- diagnostics

def _lattice_of(ch, fill):
    lat = np.zeros(ch.state_shape())
    lat[1:-1, 1:-1] = fill
    return lat


def test_diag_quiescent_zero_force(orc):
    ch = Channel(40, 40, 20, 20, 5)
    lat = _lattice_of(ch, W)
    fx, fy, ke, mass = orc.diag(ch, lat)
    assert abs(fx) < 1e-14 and abs(fy) < 1e-14 and ke == 0.0
    nsolid = sum(1 for x in range(2, 40) for y in range(2, 40)
                 if (x - 20) ** 2 + (y - 20) ** 2 <= 25)
    assert mass == pytest.approx(40 * 40 - nsolid, rel=1e-14)


def test_diag_single_link_by_hand(orc):
    """One solid node: the momentum exchanged is sum over links of
    c_i (f into the solid + f out of it) -- checked link by link."""
    ch = Channel(9, 9, 5, 5, 0)
    lat = _lattice_of(ch, 0.0)
    lat[5, 5, 1] = 0.25    # arrived in the solid from (4,5), moving +x
    lat[4, 5, 3] = 0.5     # arrived at (4,5) from the solid, moving -x
    lat[6, 6, 5] = 0.125   # arrived at (6,6) from the solid, moving +x+y
    fx, fy, _, _ = orc.diag(ch, lat)
    assert fx == pytest.approx(0.25 + 0.5 - 0.125, abs=0)
    assert fy == pytest.approx(-0.125, abs=0)


def test_diag_uniform_flow_kinetic_energy(orc):
    ch = Channel(30, 20, 0, 0, -1)
    rho, ux, uy = 1.02, 0.03, -0.01
    f = np.array([orc.equilibrium(i, rho, ux, uy, ux * ux + uy * uy) for i in range(9)])
    lat = _lattice_of(ch, f)
    _, _, ke, mass = orc.diag(ch, lat)
    assert ke == pytest.approx(30 * 20 * 0.5 * rho * (ux * ux + uy * uy), rel=1e-12)
    assert mass == pytest.approx(30 * 20 * rho, rel=1e-13)
