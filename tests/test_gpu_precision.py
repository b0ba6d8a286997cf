"""Reduced-precision storage (deviations f_i - w_i; FP32 storage: FP64
collisions rounded stochastically; FP16 storage: FP32 collisions rounded to
nearest; FP64 wall operators and reductions) against the bit-exact FP64 path.
No reference counterpart: the FP32 bound is 1e-5 relative velocity error
(L2 over both components), the FP16 bounds are the measured behaviour of a
format whose quantum is ~5e-4 relative."""
import numpy as np
import pytest

import lbm_amd as L

pytestmark = pytest.mark.gpu


def run(ch, steps, storage, variant="combine"):
    with L.Lattice(ch, storage=storage) as lat:
        lat.init_equilibrium()
        lat.step(steps, variant)
        return lat.macros(), lat.diagnostics()


def rel_l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def c1_fields():
    ch = L.Channel.geometry(256, 128, "B")
    out = {s: run(ch, 2000, s) for s in ("f64", "f32", "f16")}
    return ch, out


def test_f32_velocity_error(c1_fields):
    """Measured on MI355X: 8.3e-7 after 2000 steps.  (Before the equilibrium
    weights became two-term FP32 products the error grew linearly, 1.07e-5 at
    2000 steps and 8e-5 at 20000: FP32 1/9 and 1/36 carry the same relative
    error, 7.45e-9, which added momentum at every collision.)"""
    ch, out = c1_fields
    (r64, u64, v64), _ = out["f64"]
    (r32, u32, v32), _ = out["f32"]
    vel64 = np.stack([u64, v64]); vel32 = np.stack([u32, v32])
    err = rel_l2(vel32, vel64)
    assert err <= 1e-5, err
    assert np.abs(r32 - r64).max() <= 5e-6


def test_f32_velocity_error_1000_steps():
    """1e-5 relative at 1000 steps of the parity case (measured 8.7e-7)."""
    ch = L.Channel.geometry(256, 128, "B")
    (r64, u64, v64), _ = run(ch, 1000, "f64")
    (r32, u32, v32), _ = run(ch, 1000, "f32")
    err = rel_l2(np.stack([u32, v32]), np.stack([u64, v64]))
    assert err <= 1e-5, err


def test_f32_velocity_error_long_run():
    """20000 steps of a 1024 x 256 channel (Re 100) with the panel sweep: the
    FP32 error stays at the rounding level instead of drifting (measured
    2.1e-6; the error of the one-term weights was 8.7e-5 here)."""
    ch = L.Channel.geometry(1024, 256, "B")
    (r64, u64, v64), _ = run(ch, 20000, "f64", "panel")
    (r32, u32, v32), _ = run(ch, 20000, "f32", "panel")
    err = rel_l2(np.stack([u32, v32]), np.stack([u64, v64]))
    assert err <= 1e-5, err


def test_f16_velocity_error(c1_fields):
    ch, out = c1_fields
    (r64, u64, v64), _ = out["f64"]
    (r16, u16, v16), _ = out["f16"]
    err = rel_l2(np.stack([u16, v16]), np.stack([u64, v64]))
    # measured 1.0e-2 after 2000 steps: the FP16 storage quantum (~5e-4
    # relative) is frozen into a nearly steady flow; stochastic rounding was
    # tried and did not lower it
    assert np.isfinite(err) and err <= 3e-2, err


def test_reduced_drag_lift_close(c1_fields):
    ch, out = c1_fields
    d64 = out["f64"][1]
    d32 = out["f32"][1]
    assert abs(d32[0] - d64[0]) <= 1e-4 * abs(d64[0]) + 1e-9
    assert abs(d32[3] - d64[3]) <= 1e-6 * d64[3]


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_reduced_variants_identical(storage):
    """origin (collide + stream kernels) and combine (fused pull) do the same
    arithmetic on the same stored values, so they agree bit for bit."""
    ch = L.Channel.geometry(96, 48, "A")
    with L.Lattice(ch, storage=storage) as a, L.Lattice(ch, storage=storage) as b:
        a.init_equilibrium(); b.init_equilibrium()
        a.step(37, "origin"); b.step(37, "combine")
        la, ta, pa = a.download(); lb, tb, pb = b.download()
    assert np.array_equal(la, lb) and np.array_equal(ta, tb) and np.array_equal(pa, pb)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-7), ("f16", 1e-3)])
def test_reduced_upload_download_round_trip(orc, storage, tol):
    from oracle_lib import Channel as OC
    ch = OC.geometry(64, 32, "B")
    _, lref, tref, pref = orc.run(ch, 11)
    dch = L.Channel(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r)
    with L.Lattice(dch, storage=storage) as lat:
        lat.upload(lref, tref, pref)
        lo, to, po = lat.download()
    assert np.abs(lo[1:-1, 1:-1] - lref[1:-1, 1:-1]).max() <= tol
    assert np.abs(to[1:-1, 1:-1] - tref[1:-1, 1:-1]).max() <= tol


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("shape", [(97, 41, 0, 0), (300, 140, 5, 30), (64, 200, 33, 126),
                                   (5, 7, 3, 0), (300, 260, 69, 62), (301, 260, 20, 0)])
def test_tight_block_reduced_equals_single_steps(storage, shape):
    """Two steps per launch with the first step in LDS tiles: the first
    step's values are rounded to the storage format before they go to LDS, so
    the result equals single steps bit for bit, for every tile shape (tile_x
    columns x tile_y rows; 0 = default)."""
    lx, ly, tx, ty = shape
    ch = L.Channel(lx, ly, lx // 5, ly // 2, max(ly // 10, 1))
    with L.Lattice(ch, storage=storage) as a, L.Lattice(ch, storage=storage) as b:
        b.set_tile(tx, ty)
        a.init_equilibrium(); b.init_equilibrium()
        a.step(31, "combine"); b.step(31, "tight_block")
        la, ta, pa = a.download(); lb, tb, pb = b.download()
    assert np.array_equal(la, lb) and np.array_equal(ta, tb) and np.array_equal(pa, pb)


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("shape", [(97, 41, 0, 0), (300, 140, 5, 0), (64, 200, 33, 0),
                                   (5, 7, 3, 0), (300, 260, 16, 8), (301, 260, 17, 16)])
def test_panel_reduced_equals_single_steps(storage, shape):
    """Two steps per launch with the first step in wave registers give the
    same bits as two single steps, for reduced storage too (the first step's
    values are rounded to the storage format in registers)."""
    lx, ly, cxk, wpg = shape
    ch = L.Channel(lx, ly, lx // 5, ly // 2, max(ly // 10, 1))
    with L.Lattice(ch, storage=storage) as a, L.Lattice(ch, storage=storage) as b:
        b.set_tile(cxk, wpg)
        a.init_equilibrium(); b.init_equilibrium()
        a.step(31, "combine"); b.step(31, "panel")
        la, ta, pa = a.download(); lb, tb, pb = b.download()
    assert np.array_equal(la, lb) and np.array_equal(ta, tb) and np.array_equal(pa, pb)
