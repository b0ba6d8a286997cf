"""FP64 parity of the HIP path (through the C-ABI) with the oracle and with the
golden vectors written by the real reference.  The bar is bit-equality; the
north-star tolerance (max-abs <= 1e-12 on f_i) is asserted as well so a
failure reports by how much."""
import json
import os

import numpy as np
import pytest

import lbm_amd as L
from oracle_lib import Channel as OChannel

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-12


def dev_channel(c: OChannel) -> L.Channel:
    return L.Channel(c.lx, c.ly, c.obst_x, c.obst_y, c.obst_r, c.u_max, c.reynolds,
                     c.bc_family, c.zgb)


def inner(a):
    return a[1:-1, 1:-1, :]


def assert_same(got, want, what):
    d = float(np.abs(inner(got) - inner(want)).max())
    assert d <= TOL, f"{what}: max-abs {d:g} > {TOL:g}"
    assert np.array_equal(inner(got), inner(want)), f"{what}: not bit-identical (max-abs {d:g})"


@pytest.mark.parametrize("which", range(18))
def test_node_operator_kat(which):
    kat = np.load(os.path.join(GOLD, "kat.npz"))
    if which == 17:
        f = kat["bgk_in"]
        om = kat["bgk_omega"]
        got = np.stack([L.apply_node_op(17, f[k:k + 1], np.zeros(1), np.zeros(1), om[k])[0]
                        for k in range(len(f))])
        assert np.array_equal(got, kat["bgk_out"])
        return
    f = kat["bc_in"][which]
    p = kat["bc_p"][which]
    om = kat["bc_omega"][which]
    got = np.stack([L.apply_node_op(which, f[k:k + 1], p[k:k + 1, 0], p[k:k + 1, 1], om[k])[0]
                    for k in range(len(f))])
    want = kat["bc_out"][which]
    assert np.abs(got - want).max() <= TOL
    assert np.array_equal(got, want), str(kat["bc_names"][which])


def test_bgk_velocity_division_is_ieee(orc):
    """The device forms ux and uy with one reciprocal of rho shared by both
    quotients; each must still be the correctly rounded IEEE quotient the
    reference divides out (lb.c:134-141).  BGK outputs of the oracle (C
    division) and the device must agree bit for bit on states built to probe
    the division: zero and one-ulp numerators of both signs, rho from 1e-3 to
    1e3, populations scaled down to 2^-880 and up to 2^880, and large
    velocities."""
    rng = np.random.default_rng(7)
    base = rng.uniform(0.01, 0.3, size=(600, 9))
    cases = [base]
    z = base.copy()
    z[:, 3] = z[:, 1]; z[:, 6] = z[:, 5]; z[:, 7] = z[:, 8]      # ux numerator 0
    cases.append(z)
    z2 = z.copy()
    z2[:, 4] = z2[:, 2]; z2[:, 7] = z2[:, 5]; z2[:, 8] = z2[:, 6]  # both numerators 0
    z2[:, 3] = z2[:, 1]
    cases.append(z2)
    for sgn in (1.0, -1.0):                                       # one-ulp numerators
        u = z2.copy()
        u[:, 1] = np.nextafter(u[:, 3], u[:, 3] + sgn)
        u[:, 2] = np.nextafter(u[:, 4], u[:, 4] - sgn)
        cases.append(u)
    for scale in (1e-3, 1e-1, 1e1, 1e3, 2.0 ** -880, 2.0 ** 880):
        cases.append(base[:200] * scale)
    fast = base[:300].copy()
    fast[:, [1, 5, 8]] *= 3.0                                      # strong x flow
    fast[:, [4, 7]] *= 0.2
    cases.append(fast)
    f = np.concatenate(cases)
    for omega in (0.6, 1.0, 1.9):
        got = L.apply_node_op(17, f, np.zeros(len(f)), np.zeros(len(f)), omega)
        want = np.stack([orc.bgk(v, omega) for v in f])
        bad = np.flatnonzero(~np.all(got.view(np.int64) == want.view(np.int64), axis=1))
        assert bad.size == 0, f"omega {omega}: {bad.size} nodes differ, first {f[bad[0]]}"


@pytest.mark.parametrize("name", ["traj_32x16_reg.npz", "traj_32x16_zh.npz"])
@pytest.mark.parametrize("variant", ["origin", "combine", "tight", "tight_block", "panel"])
def test_trajectory_golden(name, variant):
    g = np.load(os.path.join(GOLD, name))
    ch = OChannel(**json.loads(str(g["channel"])))
    for s in (1, 2, 3, 10, 200):
        with L.Lattice(dev_channel(ch)) as lat:
            lat.init_equilibrium()
            lat.step(s, variant)
            lo, to, po = lat.download()
        assert_same(lo, g[f"lat_{s}"], f"{name} {variant} lattice after {s}")
        assert_same(to, g[f"tmp_{s}"], f"{name} {variant} tmpLattice after {s}")
        assert np.array_equal(po[2:ch.ly], g[f"prho_{s}"][2:ch.ly])


@pytest.mark.parametrize("geom", ["A", "B"])
@pytest.mark.parametrize("variant", ["origin", "combine", "tight", "tight_block", "panel",
                                     "panel4", "panel6"])
def test_c1_parity_2000_steps(orc, geom, variant):
    """256x128, Re 100, 2000 steps: the parity configuration, every sweep.
    The lattice's 126 interior rows are too few for the four- and six-step
    marches (they need 124 rows inside an H-row halo on each side), so
    panel4 and panel6 run as the two-step panel here, as documented;
    test_deep_sweeps_2000_steps runs them deep for as long."""
    ch = OChannel.geometry(256, 128, geom)
    ran, lat_ref, tmp_ref, prho_ref = orc.run(ch, 2000, "origin")
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        lat.step(2000, variant)
        lo, to, po = lat.download()
        rho, ux, uy = lat.macros()
    assert_same(lo, lat_ref, f"C1-{geom} {variant} lattice")
    assert_same(to, tmp_ref, f"C1-{geom} {variant} tmpLattice")
    assert np.array_equal(po[2:128], prho_ref[2:128])
    if geom == "B":
        fl = np.load(os.path.join(GOLD, "c1_B_fields.npz"))
        assert np.array_equal(rho, fl["rho"])
        assert np.array_equal(ux, fl["ux"])
        assert np.array_equal(uy, fl["uy"])


@pytest.mark.parametrize("variant", ["tight", "panel4", "panel6"])
def test_deep_sweeps_2000_steps(orc, variant):
    """512x256, disc B, Re 100, 2000 steps against the oracle, bit for bit:
    the smallest channel shape of the C1 family on which the four- and
    six-step marches run deep (250 interior rows), so the headline sweep
    covers the parity horizon of C1 with its own kernels."""
    ch = OChannel.geometry(512, 256, "B")
    _, lat_ref, tmp_ref, prho_ref = orc.run(ch, 2000, "origin")
    with L.Lattice(dev_channel(ch)) as lat:
        lat.kernel_timing(True)
        lat.init_equilibrium()
        lat.step(2000, variant)
        lo, to, po = lat.download()
        recs = lat.kernel_times()
    assert_same(lo, lat_ref, f"512x256 {variant} lattice")
    assert_same(to, tmp_ref, f"512x256 {variant} tmpLattice")
    assert np.array_equal(po[2:256], prho_ref[2:256])
    depth = {"tight": 1, "panel4": 4, "panel6": 6}[variant]
    launches = max(r["launches"] for r in recs if r["depth"] == depth)
    assert launches >= 1999 // depth - 1, recs
    if depth > 1:
        # the interior dispatch updates the rows and columns H + 2 .. ly - H - 1
        # / lx - H - 1 once per level (H = depth - 1): each node counted once
        # although the leftmost chunk overlaps its neighbour
        h = depth - 1
        rec = next(r for r in recs if r["depth"] == depth and "k_wpanel_in" in r["kernel"])
        assert rec["nodes"] == (256 - 2 * h - 2) * (512 - 2 * h - 2), rec


def test_resume_from_uploaded_state(orc):
    ch = OChannel.geometry(64, 32, "B")
    _, l7, t7, p7 = orc.run(ch, 7)
    _, l20, t20, p20 = orc.run(ch, 20)
    with L.Lattice(dev_channel(ch)) as lat:
        lat.upload(l7, t7, p7)
        lat.step(13, "combine")
        lo, to, po = lat.download()
    assert_same(lo, l20, "resumed lattice")
    assert_same(to, t20, "resumed tmpLattice")


def test_mixed_variants_match(orc):
    """Variants can be switched between calls without changing the result."""
    ch = OChannel.geometry(48, 24, "A")
    _, lref, tref, _ = orc.run(ch, 15)
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        for n, v in ((3, "origin"), (4, "combine"), (2, "origin"), (1, "tight"), (5, "origin")):
            lat.step(n, v)
        assert lat.steps_done() == 15
        lo, to, _ = lat.download()
    assert_same(lo, lref, "mixed lattice")
    assert_same(to, tref, "mixed tmpLattice")


def test_download_is_repeatable():
    ch = OChannel.geometry(40, 20, "B")
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        lat.step(9, "combine")
        a = lat.download()
        b = lat.download()
        lat.step(1, "combine")
        c = lat.download()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def test_diagnostics_match_oracle(orc):
    ch = OChannel.geometry(128, 64, "B")
    _, lref, _, _ = orc.run(ch, 300)
    want = orc.diag(ch, lref)
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        lat.step(300, "combine")
        got = lat.diagnostics()
    for k in range(4):
        scale = max(abs(want[k]), 1e-300)
        assert abs(got[k] - want[k]) <= 1e-12 * max(scale, 1e-3), (k, got[k], want[k])


def test_panel_interior_waves(orc):
    """A lattice with several bands of interior waves: the branch-free
    interior march against the oracle, bit for bit, at chunk widths that run
    two columns of loads in flight (even widths: k_wpanel_in<double, 2, 2>,
    twelve waves per workgroup) and one (odd widths: k_wpanel_in<double, 1,
    2>, eight waves per workgroup), from the narrowest chunk up.  The panel
    sweep takes its workgroup size from the instantiation, not from the
    tile's second knob."""
    ch = OChannel.geometry(300, 260, "B")
    _, lref, tref, pref = orc.run(ch, 13)
    for cxk in (16, 17, 32, 33, 20, 3, 5):
        with L.Lattice(dev_channel(ch)) as lat:
            lat.set_tile(cxk, 0)
            lat.init_equilibrium()
            lat.step(13, "panel")
            lo, to, po = lat.download()
        assert_same(lo, lref, f"panel chunk {cxk}")
        assert_same(to, tref, f"panel chunk {cxk} tmp")
        assert np.array_equal(po[2:260], pref[2:260])


@pytest.mark.parametrize("geom", ["A", "B"])
def test_panel4_against_oracle(orc, geom):
    """Four steps per launch (three levels in registers, the perimeter frame
    through the third buffer) against the oracle, bit for bit, for step
    counts that leave every remainder (a two-step launch, a single fused
    step) and for chunk widths with one and two columns in flight, up to the
    40 columns the largest lattices use."""
    ch = OChannel.geometry(600, 400, geom)
    for steps in (4, 5, 6, 7, 13):
        _, lref, tref, pref = orc.run(ch, steps)
        for cxk in (0, 5, 8, 16, 40):
            with L.Lattice(dev_channel(ch)) as lat:
                lat.set_tile(cxk, 0)
                lat.init_equilibrium()
                lat.step(steps, "panel4")
                lo, to, po = lat.download()
            what = f"panel4 {geom} steps {steps} chunk {cxk}"
            assert_same(lo, lref, what)
            assert_same(to, tref, what + " tmp")
            assert np.array_equal(po[2:400], pref[2:400]), what


def test_panel4_long_run_and_switches(orc):
    """200 steps of panel4 mixed with the other sweeps on a lattice whose
    last band of interior rows is partial (ly - 1 not a multiple of the band
    height), against the oracle."""
    ch = OChannel.geometry(700, 333, "B")
    _, lref, tref, _ = orc.run(ch, 200)
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        for n, v in ((37, "panel4"), (3, "combine"), (80, "panel4"), (2, "panel"), (78, "panel4")):
            lat.step(n, v)
        lo, to, _ = lat.download()
    assert_same(lo, lref, "panel4 mixed")
    assert_same(to, tref, "panel4 mixed tmp")


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_panel4_reduced_equals_single_steps(storage):
    """In the reduced formats four steps per launch give the bits of four
    single steps (FP32: the same stochastic rounding per collision index)."""
    ch = L.Channel(600, 400, 120, 200, 40)
    with L.Lattice(ch, storage=storage) as a, L.Lattice(ch, storage=storage) as b:
        a.init_equilibrium(); b.init_equilibrium()
        a.step(41, "combine"); b.step(41, "panel4")
        la, ta, pa = a.download(); lb, tb, pb = b.download()
    assert np.array_equal(la, lb) and np.array_equal(ta, tb) and np.array_equal(pa, pb)


@pytest.mark.parametrize("geom", ["A", "B"])
def test_panel6_against_oracle(orc, geom):
    """Six steps per launch (five levels in registers, the perimeter frame
    through two further buffers) against the oracle, bit for bit, for step
    counts that leave every remainder (a four-step, a two-step launch, a
    single fused step; 9 and 10 steps run one eight-step launch, seven levels
    in registers and the frame through three further buffers, in place of a
    six- and a two-step one) and for chunk widths with one and two columns in
    flight, up to the 64 columns the largest lattices use."""
    ch = OChannel.geometry(600, 400, geom)
    for steps in (6, 7, 8, 9, 10, 13, 19):
        _, lref, tref, pref = orc.run(ch, steps)
        for cxk in (0, 5, 8, 16, 64):
            with L.Lattice(dev_channel(ch)) as lat:
                lat.set_tile(cxk, 0)
                lat.init_equilibrium()
                lat.step(steps, "panel6")
                lo, to, po = lat.download()
            what = f"panel6 {geom} steps {steps} chunk {cxk}"
            assert_same(lo, lref, what)
            assert_same(to, tref, what + " tmp")
            assert np.array_equal(po[2:400], pref[2:400]), what


def test_panel6_launch_plan():
    """Twenty steps of the six-step sweep from a collided state run as
    6 + 6 + 8 (a two-step launch costs twice a four-step one per step), and
    21 as 6 + 6 + 8 + 1; on a lattice with an interior deep enough for six
    levels but not for eight (300 x 136), 20 run as 6 + 6 + 4 + 4."""
    big, small = L.Channel(600, 400, 120, 200, 40), L.Channel(300, 136, 60, 68, 13)
    for ch, n, want in ((big, 20, {6: 2, 8: 1}), (big, 21, {6: 2, 8: 1, 1: 1}),
                        (big, 18, {6: 3}), (big, 22, {6: 3, 4: 1}), (small, 20, {6: 2, 4: 2})):
        with L.Lattice(ch) as lat:
            lat.init_equilibrium()
            lat.step(1, "panel6")
            lat.kernel_timing(True)
            lat.step(n, "panel6")
            recs = lat.kernel_times()
        got = {}
        for r in recs:
            got[r["depth"]] = got.get(r["depth"], 0) + r["launches"]
        assert got == want, (n, recs)


@pytest.mark.parametrize("variant,lx,ly", [
    ("panel6", 300, 136),   # 124 interior rows: the six-step march, thin frame
    ("panel6", 300, 135),   # one row fewer: runs as panel4 (127 rows)
    ("panel4", 300, 132),   # 124 interior rows of the four-step march
    ("panel4", 300, 131),   # runs as panel, thin two-row strips
    ("panel", 300, 66),     # 62 interior rows: the two-step march, thin frame
    ("panel", 300, 65),     # 61: every wave on the perimeter
    ("panel", 40, 5),
])
def test_thin_frame_thresholds(orc, variant, lx, ly):
    """The interior rows of the panel sweeps run from H + 2 to ly - H (H = 1,
    3, 5 halo rows) and the frame keeps H + 1 rows along each wall, packed
    several chunks per perimeter wave; lattices at and just below the row
    counts where a sweep switches to its thin frame or falls back, bit for
    bit against the oracle, at chunk widths that leave a partial first
    chunk and at the default."""
    ch = OChannel.geometry(lx, ly, "B")
    for steps in (7, 13):
        _, lref, tref, pref = orc.run(ch, steps)
        for cxk in (0, 7, 16):
            with L.Lattice(dev_channel(ch)) as lat:
                lat.set_tile(cxk, 0)
                lat.init_equilibrium()
                lat.step(steps, variant)
                lo, to, po = lat.download()
            what = f"{variant} {lx}x{ly} steps {steps} chunk {cxk}"
            assert_same(lo, lref, what)
            assert_same(to, tref, what + " tmp")
            assert np.array_equal(po[2:ly], pref[2:ly]), what


def test_panel6_long_run_and_switches(orc):
    """200 steps of panel6 mixed with the other sweeps on a lattice whose
    last band of interior rows is partial, against the oracle; and a lattice
    too small for six levels (it runs as panel4 / panel) on the way."""
    ch = OChannel.geometry(700, 333, "B")
    _, lref, tref, _ = orc.run(ch, 200)
    with L.Lattice(dev_channel(ch)) as lat:
        lat.init_equilibrium()
        for n, v in ((37, "panel6"), (3, "combine"), (80, "panel6"), (2, "panel4"), (78, "panel6")):
            lat.step(n, v)
        lo, to, _ = lat.download()
    assert_same(lo, lref, "panel6 mixed")
    assert_same(to, tref, "panel6 mixed tmp")
    small = OChannel.geometry(90, 70, "A")
    _, lref, tref, _ = orc.run(small, 15)
    with L.Lattice(dev_channel(small)) as lat:
        lat.init_equilibrium()
        lat.step(15, "panel6")
        lo, to, _ = lat.download()
    assert_same(lo, lref, "panel6 small lattice")
    assert_same(to, tref, "panel6 small lattice tmp")


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_panel6_reduced_equals_single_steps(storage):
    """In the reduced formats six (and, for remainders of 2 or 3, eight)
    steps per launch give the bits of as many single steps (FP32: the same
    stochastic rounding per collision index): 43 = 1 + 7 x 6, 21 = 1 + 6 + 6
    + 8, 10 = 1 + 8 + 1."""
    ch = L.Channel(600, 400, 120, 200, 40)
    for n in (43, 21, 10):
        with L.Lattice(ch, storage=storage) as a, L.Lattice(ch, storage=storage) as b:
            a.init_equilibrium(); b.init_equilibrium()
            a.step(n, "combine"); b.step(n, "panel6")
            la, ta, pa = a.download(); lb, tb, pb = b.download()
        assert np.array_equal(la, lb) and np.array_equal(ta, tb) and np.array_equal(pa, pb), n


@pytest.mark.parametrize("lx,ly", [(3, 3), (3, 8), (4, 3), (5, 4), (9, 3)])
def test_smallest_lattices(orc, lx, ly):
    """The smallest channels the API accepts (3 x 3 and up), a disc on the
    corner included: every sweep bit-identical to the oracle, except that the
    two-step and deeper sweeps refuse lx < 4 with an error, as documented
    (the deeper sweeps fall back to the two-step one on lattices this small)."""
    ch = OChannel.geometry(lx, ly, "A")
    _, lref, tref, pref = orc.run(ch, 7)
    for v in ("origin", "combine", "tight", "tight_block", "panel", "panel4", "panel6"):
        with L.Lattice(dev_channel(ch)) as lat:
            lat.init_equilibrium()
            if lx < 4 and v in ("tight_block", "panel", "panel4", "panel6"):
                with pytest.raises(L.LbmError):
                    lat.step(7, v)
                continue
            lat.step(7, v)
            lo, to, po = lat.download()
        assert_same(lo, lref, f"{v} {lx}x{ly}")
        assert_same(to, tref, f"{v} {lx}x{ly} tmp")
        assert np.array_equal(po[2:ly], pref[2:ly])


@pytest.mark.parametrize("variant,tiles", [
    ("panel", [(0, 0), (3, 5), (7, 254), (130, 17), (4, 1)]),
    ("tight_block", [(0, 0), (3, 30), (5, 62), (20, 126), (50, 62)]),
])
def test_two_step_tile_shapes(orc, variant, tiles):
    """Bit-identical for every tile / panel shape, including tiles that do not
    divide the lattice and lattices smaller than one tile."""
    for lx, ly, geom in ((61, 37, "B"), (23, 100, "A"), (4, 9, "A")):
        ch = OChannel.geometry(lx, ly, geom)
        _, lref, tref, pref = orc.run(ch, 13)
        for tx, ty in tiles:
            with L.Lattice(dev_channel(ch)) as lat:
                lat.set_tile(tx, ty)
                lat.init_equilibrium()
                lat.step(13, variant)
                lo, to, po = lat.download()
            assert_same(lo, lref, f"{variant} {lx}x{ly} tile {tx}x{ty}")
            assert_same(to, tref, f"{variant} {lx}x{ly} tile {tx}x{ty} tmp")
            assert np.array_equal(po[2:ly], pref[2:ly])


@pytest.mark.parametrize("variant", ["combine", "panel"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("graph_steps", [6, 7])
def test_graph_replay_equals_eager(variant, storage, graph_steps):
    """hipGraph replays (fused steps, or graph_steps / 2 panel launches over
    two streams) give the same bits as eager launches, replay after replay,
    with eager steps and downloads in between; for FP32 this also checks that
    the device collision counter advances inside replays."""
    ch = L.Channel(300, 260, 60, 130, 26)
    with L.Lattice(ch, storage=storage) as eager, L.Lattice(ch, storage=storage) as graph:
        graph.set_graph_steps(graph_steps)
        for lat in (eager, graph):
            lat.init_equilibrium()
        for chunk in (1, 2 * graph_steps, 3, graph_steps, 5 * graph_steps + 1, 2 * graph_steps):
            eager.step(chunk, variant)
            graph.step(chunk, variant)
            le, te, pe = eager.download()
            lg, tg, pg = graph.download()
            assert np.array_equal(le, lg) and np.array_equal(te, tg), (chunk, variant, storage)
            assert np.array_equal(pe, pg)
        assert eager.steps_done() == graph.steps_done()
