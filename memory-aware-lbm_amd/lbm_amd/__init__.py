"""Python binding of the MI355X D2Q9 collide-stream library (liblbm_hip.so).

A thin ctypes layer over the C-ABI of include/lbm_hip.h, used by the tests,
bench.py and the batching/serving drivers.  It has no compute of its own: if
the HIP library is missing every entry point raises, so nothing can silently
fall back to the CPU.

The channel helpers build the reference driver's problem
(IUPU-HPC/memory-aware-LBM unsteady.c:84-199): bulk BGK, bounce-back disc,
regularized velocity walls (u = 0), a Poiseuille inlet and a pressure outlet
with zero-gradient density extrapolation.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "liblbm_hip.so")

ORIGIN, COMBINE, TIGHT, TIGHT_BLOCK, PANEL, PANEL4, PANEL6 = range(7)
VARIANT = {"origin": ORIGIN, "combine": COMBINE, "tight": TIGHT,
           "tight_block": TIGHT_BLOCK, "panel": PANEL, "panel4": PANEL4,
           "panel6": PANEL6}
STORE_F64, STORE_F32, STORE_F16 = range(3)
STORAGE = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16}
BC_REGULARIZED, BC_ZOUHE = 0, 1
BC_VELOCITY, BC_PRESSURE = 0, 1
UPPER, LOWER, LEFT, RIGHT = range(4)

_P = C.POINTER(C.c_double)


class LbmError(RuntimeError):
    pass


class _SideBC(C.Structure):
    _fields_ = [("kind", C.c_int), ("family", C.c_int), ("p0", _P), ("p1", _P)]


class _KernelRecord(C.Structure):
    _fields_ = [("kernel", C.c_char * 64), ("depth", C.c_int), ("launches", C.c_int),
                ("ms", C.c_double), ("min_ms", C.c_double), ("max_ms", C.c_double),
                ("nodes", C.c_double)]


class _Problem(C.Structure):
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("obst_x", C.c_int),
                ("obst_y", C.c_int), ("obst_r", C.c_int), ("omega", C.c_double),
                ("side", _SideBC * 4), ("zgb", C.c_int), ("storage", C.c_int)]


_lib = None


def library():
    """Load liblbm_hip.so; raises LbmError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LbmError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "(make -C memory-aware-lbm_amd)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "lbm_hip_create": (C.c_int, [C.POINTER(_Problem), C.POINTER(vp)]),
        "lbm_hip_destroy": (C.c_int, [vp]),
        "lbm_hip_init_equilibrium": (C.c_int, [vp, _P]),
        "lbm_hip_upload": (C.c_int, [vp, _P, _P, _P]),
        "lbm_hip_download": (C.c_int, [vp, _P, _P, _P]),
        "lbm_hip_step": (C.c_int, [vp, C.c_int, C.c_int]),
        "lbm_hip_set_tile": (C.c_int, [vp, C.c_int, C.c_int]),
        "lbm_hip_set_side_params": (C.c_int, [vp, C.c_int, _P, _P]),
        "lbm_hip_macros": (C.c_int, [vp, _P, _P, _P]),
        "lbm_hip_diagnostics": (C.c_int, [vp, _P]),
        "lbm_hip_steps_done": (C.c_int64, [vp]),
        "lbm_hip_sync": (C.c_int, [vp]),
        "lbm_hip_time_steps": (C.c_int, [vp, C.c_int, C.c_int, _P]),
        "lbm_hip_bytes_per_node": (C.c_double, [vp, C.c_int]),
        "lbm_hip_footprint": (C.c_size_t, [vp]),
        "lbm_hip_footprint_of": (C.c_size_t, [C.POINTER(_Problem)]),
        "lbm_hip_footprint_for": (C.c_size_t, [C.POINTER(_Problem), C.c_int]),
        "lbm_hip_apply_node_op": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, C.c_double]),
        "lbm_hip_copy_bandwidth": (C.c_int, [C.c_size_t, C.c_int, _P]),
        "lbm_hip_set_graph_steps": (C.c_int, [vp, C.c_int]),
        "lbm_hip_make_periodic": (C.c_int, [vp]),
        "lbm_hip_query": (C.c_int, [vp]),
        "lbm_hip_device_memory": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "lbm_hip_pool_reserve": (C.c_int, [C.c_size_t]),
        "lbm_hip_pool_release": (C.c_int, []),
        "lbm_hip_kernel_timing": (C.c_int, [vp, C.c_int]),
        "lbm_hip_kernel_times": (C.c_int, [vp, C.POINTER(_KernelRecord), C.c_int,
                                           C.POINTER(C.c_int)]),
        "lbm_hip_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = library().lbm_hip_last_error().decode(errors="replace")
        raise LbmError(f"{what} failed ({rc}): {msg}")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_P)


@dataclass
class Channel:
    """Cylinder channel of the reference driver with an explicit disc."""

    lx: int
    ly: int
    obst_x: int
    obst_y: int
    obst_r: int
    u_max: float = 0.02
    reynolds: float = 100.0
    bc_family: int = BC_REGULARIZED
    zgb: int = 1

    @staticmethod
    def geometry(lx: int, ly: int, geom: str = "B", **kw) -> "Channel":
        """"A": the reference driver's disc (lx/4, ly/2, ly/10+1),
        unsteady.c:96-98; "B": (lx/5, ly/2, ly/10)."""
        if geom == "A":
            return Channel(lx, ly, lx // 4, ly // 2, ly // 10 + 1, **kw)
        if geom == "B":
            return Channel(lx, ly, lx // 5, ly // 2, ly // 10, **kw)
        raise ValueError(geom)

    def omega(self) -> float:
        nu = self.u_max * 2.0 * self.obst_r / self.reynolds  # unsteady.c:102
        return 1.0 / (3 * nu + 1.0 / 2.0)                    # unsteady.c:103

    def poiseuille(self) -> np.ndarray:
        """Inlet / initial velocity per row y = 0..ly+1 (unsteady.c:155-159)."""
        out = np.zeros(self.ly + 2)
        L = float(self.ly - 1)
        for y in range(self.ly + 2):
            yy = float(y - 1)
            out[y] = 4.0 * self.u_max / (L * L) * (L * yy - yy * yy)
        return out


class Lattice:
    """One device context (lbm_ctx) holding a channel's populations."""

    def __init__(self, ch: Channel, storage: str = "f64"):
        lib = library()
        self.ch = ch
        self.storage = storage
        u = ch.poiseuille()
        inlet_p0 = u.copy()
        inlet_p0[0] = inlet_p0[1] = inlet_p0[ch.ly] = inlet_p0[ch.ly + 1] = 0.0
        # the reference's inlet data exists for rows 2..ly-1 only
        self._keep = []
        zeros_x = np.zeros(ch.lx + 2)
        zeros_y = np.zeros(ch.ly + 2)
        outlet = np.zeros(ch.ly + 2)
        outlet[2:ch.ly] = 1.0
        self._keep += [inlet_p0, zeros_x, zeros_y, outlet]
        prob = _Problem()
        prob.lx, prob.ly = ch.lx, ch.ly
        prob.obst_x, prob.obst_y, prob.obst_r = ch.obst_x, ch.obst_y, ch.obst_r
        prob.omega = ch.omega()
        fam = ch.bc_family
        prob.side[UPPER] = _SideBC(BC_VELOCITY, fam, _ptr(zeros_x), _ptr(zeros_x))
        prob.side[LOWER] = _SideBC(BC_VELOCITY, fam, _ptr(zeros_x), _ptr(zeros_x))
        prob.side[LEFT] = _SideBC(BC_VELOCITY, fam, _ptr(inlet_p0), _ptr(zeros_y))
        prob.side[RIGHT] = _SideBC(BC_PRESSURE, fam, _ptr(outlet), _ptr(zeros_y))
        prob.zgb = ch.zgb
        prob.storage = STORAGE[storage]
        self._prob = prob
        h = C.c_void_p()
        _check(lib.lbm_hip_create(C.byref(prob), C.byref(h)), "lbm_hip_create")
        self._h = h
        self._init_u = u

    def close(self):
        if getattr(self, "_h", None):
            library().lbm_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state ---------------------------------------------------------------
    def init_equilibrium(self):
        _check(library().lbm_hip_init_equilibrium(self._h, _ptr(self._init_u)), "init")
        return self

    def upload(self, lattice: np.ndarray, tmp: np.ndarray, outlet_rho=None):
        lat = np.ascontiguousarray(lattice, dtype=np.float64)
        tm = np.ascontiguousarray(tmp, dtype=np.float64)
        pr = None if outlet_rho is None else np.ascontiguousarray(outlet_rho, dtype=np.float64)
        _check(library().lbm_hip_upload(self._h, _ptr(lat), _ptr(tm), _ptr(pr)), "upload")

    def download(self):
        """(lattice, tmpLattice, outlet_rho) in the reference's convention;
        ghost rings are zero."""
        shape = (self.ch.lx + 2, self.ch.ly + 2, 9)
        lat = np.zeros(shape)
        tmp = np.zeros(shape)
        pr = np.zeros(self.ch.ly + 2)
        _check(library().lbm_hip_download(self._h, _ptr(lat), _ptr(tmp), _ptr(pr)), "download")
        return lat, tmp, pr

    def step(self, n: int = 1, variant: str = "combine"):
        _check(library().lbm_hip_step(self._h, VARIANT[variant], int(n)), "step")

    def make_periodic(self):
        """makePeriodic (lb.c:1783-1817) on the post-stream state."""
        _check(library().lbm_hip_make_periodic(self._h), "make_periodic")

    def set_tile(self, tx: int, ty: int):
        _check(library().lbm_hip_set_tile(self._h, tx, ty), "set_tile")

    def sync(self):
        _check(library().lbm_hip_sync(self._h), "sync")

    def set_graph_steps(self, k: int):
        _check(library().lbm_hip_set_graph_steps(self._h, int(k)), "set_graph_steps")

    def done(self) -> bool:
        """True when all enqueued work has finished (non-blocking)."""
        rc = library().lbm_hip_query(self._h)
        if rc < 0:
            _check(rc, "query")
        return rc == 1

    def steps_done(self) -> int:
        return int(library().lbm_hip_steps_done(self._h))

    def macros(self):
        n = self.ch.lx * self.ch.ly
        r, u, v = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(library().lbm_hip_macros(self._h, _ptr(r), _ptr(u), _ptr(v)), "macros")
        s = (self.ch.lx, self.ch.ly)
        return r.reshape(s), u.reshape(s), v.reshape(s)

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(4)
        _check(library().lbm_hip_diagnostics(self._h, _ptr(out)), "diagnostics")
        return out

    def time_steps(self, n: int, variant: str = "combine") -> float:
        ms = C.c_double()
        _check(library().lbm_hip_time_steps(self._h, VARIANT[variant], int(n),
                                            C.byref(ms)), "time_steps")
        return ms.value

    def kernel_timing(self, on: bool = True):
        """Start (or stop) recording the launches of the kernels that carry
        the lattice updates; clears the record."""
        _check(library().lbm_hip_kernel_timing(self._h, int(bool(on))), "kernel_timing")

    def kernel_times(self) -> list:
        """One dict per (kernel, steps per launch) recorded since
        kernel_timing(True): kernel, depth, launches, ms (summed), avg_ms,
        min_ms, max_ms, nodes (updated per launch)."""
        n = C.c_int()
        _check(library().lbm_hip_kernel_times(self._h, None, 0, C.byref(n)), "kernel_times")
        recs = (_KernelRecord * max(n.value, 1))()
        _check(library().lbm_hip_kernel_times(self._h, recs, n.value, C.byref(n)), "kernel_times")
        out = []
        for r in recs[:n.value]:
            out.append({"kernel": r.kernel.decode(), "depth": r.depth, "launches": r.launches,
                        "ms": r.ms, "avg_ms": r.ms / r.launches if r.launches else 0.0,
                        "min_ms": r.min_ms, "max_ms": r.max_ms, "nodes": r.nodes})
        return out

    def bytes_per_node(self, variant: str) -> float:
        return float(library().lbm_hip_bytes_per_node(self._h, VARIANT[variant]))

    def footprint(self) -> int:
        return int(library().lbm_hip_footprint(self._h))


def device_memory():
    """(free, total) device bytes."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(library().lbm_hip_device_memory(C.byref(f), C.byref(t)), "device_memory")
    return int(f.value), int(t.value)


def pool_reserve(nbytes: int = 0):
    """Give the current device a stream-ordered memory pool for contexts
    (grown by nbytes up front): contexts then come and go without waiting
    for other contexts' work."""
    _check(library().lbm_hip_pool_reserve(int(nbytes)), "pool_reserve")


def pool_release():
    """Destroy the device's context pool (all its contexts must be gone)."""
    _check(library().lbm_hip_pool_release(), "pool_release")


def footprint_of(ch: "Channel", storage: str = "f64", variant: Optional[str] = None) -> int:
    """Device bytes a context for this channel would hold (no allocation);
    with a variant, including the further buffers that sweep allocates."""
    p = _Problem()
    p.lx, p.ly, p.storage = ch.lx, ch.ly, STORAGE[storage]
    if variant is None:
        return int(library().lbm_hip_footprint_of(C.byref(p)))
    return int(library().lbm_hip_footprint_for(C.byref(p), VARIANT[variant]))


def copy_bandwidth(nbytes: int, iters: int = 20) -> float:
    """GB/s of a plain device copy kernel moving nbytes in and nbytes out."""
    ms = C.c_double()
    _check(library().lbm_hip_copy_bandwidth(int(nbytes), int(iters), C.byref(ms)), "copy")
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def apply_node_op(which: int, f: np.ndarray, p0: np.ndarray, p1: np.ndarray,
                  omega: float) -> np.ndarray:
    """Run one device node operator on n population vectors (n, 9)."""
    g = np.ascontiguousarray(f, dtype=np.float64).copy()
    n = g.shape[0]
    a = np.ascontiguousarray(p0, dtype=np.float64)
    b = np.ascontiguousarray(p1, dtype=np.float64)
    _check(library().lbm_hip_apply_node_op(int(which), n, _ptr(g), _ptr(a), _ptr(b),
                                           float(omega)), "apply_node_op")
    return g
