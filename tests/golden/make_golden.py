"""Write the golden fixtures of this directory from the REAL reference.

Runs in the development container only, where the reference checkout exists:
    make -C oracle ref && python tests/golden/make_golden.py

The reference's lb.c and boundaries.c are compiled unmodified (oracle/Makefile)
behind oracle/ref_harness.c, which plays the role of the reference driver
unsteady.c with an explicit disc geometry.  Everything written here is data:
inputs and the reference's outputs.

Fixtures
  kat.npz        per-function known answers on seeded random inputs:
                 computeMacros (lb.c:134), computeEquilibrium (lb.c:144),
                 bgk (lb.c:154) and the 17 node operators of boundaries.c.
  traj_32x16_{reg,zh}.npz
                 full lattice and tmpLattice (ghost ring included) and the
                 outlet density after 1, 2, 3, 10 and 200 steps of the
                 origin sweep, regularized and Zou/He walls.  Steps 1-3 pin
                 the stale-slot rule at the corners.
  traj_64x32_reg.npz
                 the same after 200 steps on 64x32.
  c1_{A,B}.json + c1_B_fields.npz
                 256x128, Re 100, 2000 steps (the parity configuration):
                 SHA-256 of the lattice and tmpLattice bytes, and (B) the
                 density and velocity fields by computeMacros.
  variants.json  serial origin/combine/tight/tight_block/panel bit-equality
                 on the parity lattices, and the OpenMP flavours' agreement.
  periodic.npz   makePeriodic (lb.c:1783-1817) on a seeded random 11x8
                 lattice, and a 32x16 origin trajectory of 12 steps with
                 makePeriodic applied after step 5.

`python tests/golden/make_golden.py periodic` writes only periodic.npz.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import BC_NAMES, Channel, reference  # noqa: E402

SEED = 20181
TRAJ_STEPS = (1, 2, 3, 10, 200)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def fields(ref, lat, lx, ly):
    inner = lat[1:lx + 1, 1:ly + 1, :]
    out = np.zeros((3, lx, ly))
    for x in range(lx):
        for y in range(ly):
            out[:, x, y] = ref.macros(inner[x, y])
    return out


def random_pops(rng, n):
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    return w[None, :] * (1.0 + 0.2 * rng.standard_normal((n, 9)))


def make_kat(ref):
    rng = np.random.default_rng(SEED)
    n = 64
    f = random_pops(rng, n)
    macros = np.array([ref.macros(v) for v in f])
    eq_in = np.column_stack([
        rng.integers(0, 9, n), 1.0 + 0.1 * rng.standard_normal(n),
        0.1 * rng.standard_normal(n), 0.1 * rng.standard_normal(n)])
    eq_in = np.column_stack([eq_in, eq_in[:, 2] ** 2 + eq_in[:, 3] ** 2])
    eq_out = np.array([ref.equilibrium(int(r[0]), r[1], r[2], r[3], r[4]) for r in eq_in])
    omegas = rng.uniform(0.5, 1.95, n)
    bgk_out = np.array([ref.bgk(v, w) for v, w in zip(f, omegas)])
    nb = 32
    bc_in = random_pops(rng, 17 * nb).reshape(17, nb, 9)
    bc_p = np.zeros((17, nb, 2))
    for k in range(17):
        if 5 <= k <= 8 or k >= 13:   # pressure operators: (rho, uPar)
            bc_p[k, :, 0] = 1.0 + 0.05 * rng.standard_normal(nb)
            bc_p[k, :, 1] = 0.02 * rng.standard_normal(nb)
        else:                        # velocity operators: (ux, uy)
            bc_p[k, :, :] = 0.05 * rng.standard_normal((nb, 2))
    bc_om = rng.uniform(0.5, 1.95, (17, nb))
    bc_out = np.zeros_like(bc_in)
    for k in range(17):
        for j in range(nb):
            bc_out[k, j] = ref.apply_bc(k, bc_in[k, j], bc_p[k, j, 0], bc_p[k, j, 1],
                                        bc_om[k, j])
    np.savez_compressed(
        os.path.join(HERE, "kat.npz"),
        macros_in=f, macros_out=macros, eq_in=eq_in, eq_out=eq_out,
        bgk_in=f, bgk_omega=omegas, bgk_out=bgk_out,
        bc_in=bc_in, bc_p=bc_p, bc_omega=bc_om, bc_out=bc_out,
        bc_names=np.array(BC_NAMES))


def make_traj(ref, ch: Channel, name: str, steps_list):
    arrays = {}
    for s in steps_list:
        ran, lat, tmp, prho = ref.run(ch, s, "origin")
        assert ran == s
        arrays[f"lat_{s}"] = lat
        arrays[f"tmp_{s}"] = tmp
        arrays[f"prho_{s}"] = prho
    arrays["channel"] = np.array(json.dumps(ch.__dict__))
    np.savez_compressed(os.path.join(HERE, name), **arrays)


def make_c1(ref):
    out = {}
    for geom in ("A", "B"):
        ch = Channel.geometry(256, 128, geom)
        ran, lat, tmp, prho = ref.run(ch, 2000, "origin")
        rec = dict(channel=ch.__dict__, steps=ran, omega=ch.omega(),
                   sha_lattice=sha(lat), sha_tmp=sha(tmp), sha_prho=sha(prho),
                   lattice_sum=float(lat.sum()), tmp_sum=float(tmp.sum()))
        with open(os.path.join(HERE, f"c1_{geom}.json"), "w") as fh:
            json.dump(rec, fh, indent=1)
        out[geom] = (ch, lat, tmp, prho)
        if geom == "B":
            fl = fields(ref, lat, ch.lx, ch.ly)
            np.savez_compressed(os.path.join(HERE, "c1_B_fields.npz"),
                                rho=fl[0], ux=fl[1], uy=fl[2])
    return out


def make_variants(ref, c1):
    rec = {"serial": {}, "openmp": {}}
    for geom, (ch, lat0, tmp0, _) in c1.items():
        for v in ("combine", "tight", "tight_block", "panel"):
            ran, lat, tmp, _ = ref.run(ch, 2000, v, blk=64 if v != "panel" else 32)
            rec["serial"][f"{geom}:{v}"] = bool(np.array_equal(lat, lat0) and
                                                np.array_equal(tmp, tmp0))
    # OpenMP flavours: 4 threads, thread_block 64, blk 32, on geometry B.
    ch, lat0, tmp0, _ = c1["B"]
    for v in ("origin", "combine", "tight", "tight_block", "panel"):
        ran, lat, tmp, _ = ref.run(ch, 2000, v, blk=32, openmp=1, nthreads=4, tb=64)
        rec["openmp"][f"B:{v}"] = dict(
            bitwise=bool(np.array_equal(lat[1:-1, 1:-1], lat0[1:-1, 1:-1])),
            max_abs=float(np.abs(lat[1:-1, 1:-1] - lat0[1:-1, 1:-1]).max()))
    with open(os.path.join(HERE, "variants.json"), "w") as fh:
        json.dump(rec, fh, indent=1)
    return rec


def make_periodic(ref):
    rng = np.random.default_rng(SEED + 1)
    lx, ly = 11, 8
    lat_in = rng.standard_normal((lx + 2, ly + 2, 9))
    lat_out = ref.make_periodic(lx, ly, lat_in)
    ch = Channel.geometry(32, 16, "A")
    ran, lat, tmp, prho = ref.run_periodic(ch, 12, 5, "origin")
    assert ran == 12
    np.savez_compressed(os.path.join(HERE, "periodic.npz"), kat_lx=lx, kat_ly=ly,
                        kat_in=lat_in, kat_out=lat_out, traj_steps=12, traj_after=5,
                        lat=lat, tmp=tmp, prho=prho, channel=np.array(json.dumps(ch.__dict__)))


def main():
    ref = reference()
    if not ref.available:
        sys.exit("oracle/_ref/libref_harness.so is missing: run `make -C oracle ref` "
                 "where the reference checkout exists")
    if sys.argv[1:] == ["periodic"]:
        make_periodic(ref)
        return
    make_periodic(ref)
    make_kat(ref)
    make_traj(ref, Channel.geometry(32, 16, "A"), "traj_32x16_reg.npz", TRAJ_STEPS)
    make_traj(ref, Channel.geometry(32, 16, "A", bc_family=1), "traj_32x16_zh.npz",
              TRAJ_STEPS)
    make_traj(ref, Channel.geometry(64, 32, "B"), "traj_64x32_reg.npz", (200,))
    c1 = make_c1(ref)
    rec = make_variants(ref, c1)
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
