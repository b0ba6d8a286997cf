"""ctypes access to the CPU oracle (oracle/liblbm_oracle.so) and, where it was
built, to the reference harness (oracle/_ref/libref_harness.so).

Test infrastructure: imported by tests/, tests/golden/make_golden.py,
__graft_entry__.smoke() and bench.py's cpu_baseline leg only, always as the
checker.  Both libraries export the same run/kat entry points, so the same
helpers drive either.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, asdict

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liblbm_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")

_P = C.POINTER(C.c_double)

VARIANTS = {"origin": 0, "combine": 1, "tight": 2, "tight_block": 3, "panel": 4}
BC_NAMES = [
    "bounceBack",
    "upperZouHe", "lowerZouHe", "leftZouHe", "rightZouHe",
    "upperPressureZouHe", "lowerPressureZouHe", "leftPressureZouHe", "rightPressureZouHe",
    "upperRegularized", "lowerRegularized", "leftRegularized", "rightRegularized",
    "upperPressureRegularized", "lowerPressureRegularized", "leftPressureRegularized",
    "rightPressureRegularized",
]


class _Cfg(C.Structure):
    _fields_ = [
        ("lx", C.c_int), ("ly", C.c_int),
        ("obst_x", C.c_int), ("obst_y", C.c_int), ("obst_r", C.c_int),
        ("u_max", C.c_double), ("reynolds", C.c_double),
        ("bc_family", C.c_int), ("variant", C.c_int),
        ("openmp", C.c_int), ("nthreads", C.c_int),
        ("blk", C.c_int), ("tb", C.c_int),
        ("steps", C.c_int), ("zgb", C.c_int),
    ]


@dataclass
class Channel:
    """The cylinder channel of the reference driver (unsteady.c:84-199) with an
    explicit disc.  Geometry "A" is the driver's own (lx/4, ly/2, ly/10+1);
    geometry "B" is (lx/5, ly/2, ly/10)."""

    lx: int
    ly: int
    obst_x: int
    obst_y: int
    obst_r: int
    u_max: float = 0.02
    reynolds: float = 100.0
    bc_family: int = 0  # 0 regularized, 1 Zou/He
    zgb: int = 1

    @staticmethod
    def geometry(lx: int, ly: int, geom: str = "B", **kw) -> "Channel":
        if geom == "A":
            return Channel(lx, ly, lx // 4, ly // 2, ly // 10 + 1, **kw)
        if geom == "B":
            return Channel(lx, ly, lx // 5, ly // 2, ly // 10, **kw)
        raise ValueError(geom)

    def omega(self) -> float:
        # unsteady.c:102-103, evaluated in the same order
        nu = self.u_max * 2.0 * self.obst_r / self.reynolds
        return 1.0 / (3 * nu + 1.0 / 2.0)

    def state_shape(self):
        return (self.lx + 2, self.ly + 2, 9)


def _load(path: str):
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    prefix = "orc" if "oracle.so" in path else "ref"
    run = getattr(lib, f"{prefix}_run")
    run.argtypes = [C.POINTER(_Cfg), _P, _P, _P]
    run.restype = C.c_int
    m = getattr(lib, f"{prefix}_macros")
    m.argtypes = [_P, _P]
    m.restype = None
    e = getattr(lib, f"{prefix}_equilibrium")
    e.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    e.restype = C.c_double
    b = getattr(lib, f"{prefix}_bgk")
    b.argtypes = [_P, C.c_double]
    b.restype = None
    a = getattr(lib, f"{prefix}_apply_bc")
    a.argtypes = [C.c_int, _P, C.c_double, C.c_double, C.c_double]
    a.restype = C.c_int
    mp = getattr(lib, f"{prefix}_make_periodic")
    mp.argtypes = [C.c_int, C.c_int, _P]
    mp.restype = None
    lib._prefix = prefix
    if prefix == "ref":
        lib.ref_run_periodic.argtypes = [C.POINTER(_Cfg), C.c_int, _P, _P, _P]
        lib.ref_run_periodic.restype = C.c_int
    if prefix == "orc":
        lib.orc_run_from.argtypes = [C.POINTER(_Cfg), _P, _P, _P, _P, _P, _P]
        lib.orc_run_from.restype = C.c_int
        lib.orc_diag.argtypes = [C.POINTER(_Cfg), _P, _P]
        lib.orc_diag.restype = None
    return lib


class Impl:
    """One CPU implementation (the oracle or the reference harness)."""

    def __init__(self, path: str):
        self.lib = _load(path)
        self.path = path

    @property
    def available(self) -> bool:
        return self.lib is not None

    def _fn(self, name):
        return getattr(self.lib, f"{self.lib._prefix}_{name}")

    def _cfg(self, ch: Channel, variant: str, steps: int, blk: int = 8,
             openmp: int = 0, nthreads: int = 1, tb: int = 8) -> _Cfg:
        return _Cfg(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r, ch.u_max,
                    ch.reynolds, ch.bc_family, VARIANTS[variant], openmp,
                    nthreads, blk, tb, steps, ch.zgb)

    def run(self, ch: Channel, steps: int, variant: str = "origin", blk: int = 8,
            openmp: int = 0, nthreads: int = 1, tb: int = 8):
        """Returns (steps_run, lattice, tmpLattice, outlet_rho); the lattices
        have shape (lx+2, ly+2, 9) and include the ghost ring."""
        cfg = self._cfg(ch, variant, steps, blk, openmp, nthreads, tb)
        lat = np.zeros(ch.state_shape())
        tmp = np.zeros(ch.state_shape())
        prho = np.zeros(ch.ly + 2)
        ran = self._fn("run")(C.byref(cfg), lat.ctypes.data_as(_P),
                              tmp.ctypes.data_as(_P), prho.ctypes.data_as(_P))
        if ran < 0:
            raise RuntimeError(f"{self.path}: run failed ({ran})")
        return ran, lat, tmp, prho

    def run_from(self, ch: Channel, steps: int, lat, tmp, prho, variant="origin", blk=8):
        cfg = self._cfg(ch, variant, steps, blk)
        lat_in = np.ascontiguousarray(lat, dtype=np.float64)
        tmp_in = np.ascontiguousarray(tmp, dtype=np.float64)
        prho_in = np.ascontiguousarray(prho, dtype=np.float64)
        lo = np.zeros(ch.state_shape())
        to = np.zeros(ch.state_shape())
        po = np.zeros(ch.ly + 2)
        ran = self.lib.orc_run_from(C.byref(cfg), lat_in.ctypes.data_as(_P),
                                    tmp_in.ctypes.data_as(_P), prho_in.ctypes.data_as(_P),
                                    lo.ctypes.data_as(_P), to.ctypes.data_as(_P),
                                    po.ctypes.data_as(_P))
        if ran < 0:
            raise RuntimeError("orc_run_from failed")
        return ran, lo, to, po

    def make_periodic(self, lx: int, ly: int, lat) -> np.ndarray:
        """makePeriodic (lb.c:1783-1817) on a copy of a (lx+2, ly+2, 9) lattice."""
        out = np.array(lat, dtype=np.float64, order="C")
        assert out.shape == (lx + 2, ly + 2, 9)
        self._fn("make_periodic")(int(lx), int(ly), out.ctypes.data_as(_P))
        return out

    def run_periodic(self, ch: Channel, steps: int, after: int, variant: str = "origin"):
        """steps steps with makePeriodic applied once after step `after`:
        the reference harness runs it inside its own time loop, the oracle
        as run(after) + make_periodic + run_from(steps - after)."""
        if self.lib._prefix == "ref":
            cfg = self._cfg(ch, variant, steps)
            lat = np.zeros(ch.state_shape())
            tmp = np.zeros(ch.state_shape())
            prho = np.zeros(ch.ly + 2)
            ran = self.lib.ref_run_periodic(C.byref(cfg), int(after), lat.ctypes.data_as(_P),
                                            tmp.ctypes.data_as(_P), prho.ctypes.data_as(_P))
            if ran < 0:
                raise RuntimeError("ref_run_periodic failed")
            return ran, lat, tmp, prho
        r1, lat, tmp, prho = self.run(ch, after, variant)
        lat = self.make_periodic(ch.lx, ch.ly, lat)
        r2, lat, tmp, prho = self.run_from(ch, steps - after, lat, tmp, prho, variant)
        return r1 + r2, lat, tmp, prho

    def diag(self, ch: Channel, lat) -> np.ndarray:
        cfg = self._cfg(ch, "origin", 0)
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        out = np.zeros(4)
        self.lib.orc_diag(C.byref(cfg), lat.ctypes.data_as(_P), out.ctypes.data_as(_P))
        return out

    def macros(self, f) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros(3)
        self._fn("macros")(f.ctypes.data_as(_P), out.ctypes.data_as(_P))
        return out

    def equilibrium(self, i, rho, ux, uy, usqr) -> float:
        return self._fn("equilibrium")(int(i), float(rho), float(ux), float(uy), float(usqr))

    def bgk(self, f, omega) -> np.ndarray:
        g = np.array(f, dtype=np.float64)
        self._fn("bgk")(g.ctypes.data_as(_P), float(omega))
        return g

    def apply_bc(self, which, f, p0, p1, omega) -> np.ndarray:
        g = np.array(f, dtype=np.float64)
        rc = self._fn("apply_bc")(int(which), g.ctypes.data_as(_P), float(p0), float(p1),
                                  float(omega))
        if rc:
            raise ValueError(which)
        return g


def oracle() -> Impl:
    return Impl(ORACLE_SO)


def reference() -> Impl:
    return Impl(REF_SO)


def channel_dict(ch: Channel) -> dict:
    return asdict(ch)
