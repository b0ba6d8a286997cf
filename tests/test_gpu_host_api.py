"""The reference's node-operator API on host nodes (lb.h bgk, boundaries.h's
17 operators), called through the host mirror liblb.so exactly as a
reference driver calls them (fPop, selfData): bit-identical to the known
answers the reference wrote (tests/golden/kat.npz).  Each call runs the
device operator on the one node through the library's fixed staging array."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_lib import BC_NAMES

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
_P = C.POINTER(C.c_double)


class Dynamics(C.Structure):
    _fields_ = [("dynamicsFun", C.c_void_p), ("selfData", C.c_void_p)]


class VelocityBCData(C.Structure):          # boundaries.h:50-56
    _fields_ = [("bulkDynamics", C.POINTER(Dynamics)), ("ux", C.c_double), ("uy", C.c_double)]


class PressureBCData(C.Structure):          # boundaries.h:58-64
    _fields_ = [("bulkDynamics", C.POINTER(Dynamics)), ("rho", C.c_double),
                ("uPar", C.c_double)]


@pytest.fixture(scope="module")
def liblb():
    import lbm_amd
    lbm_amd.library()                       # loads liblbm_hip.so first
    lib = C.CDLL(os.path.join(ROOT, "memory-aware-lbm_amd", "liblb.so"))
    for name in BC_NAMES + ["bgk"]:
        getattr(lib, name).argtypes = [_P, C.c_void_p]
        getattr(lib, name).restype = None
    return lib


def test_bgk_on_host_nodes(liblb):
    kat = np.load(os.path.join(GOLD, "kat.npz"))
    for f, om, want in zip(kat["bgk_in"], kat["bgk_omega"], kat["bgk_out"]):
        g = np.array(f, dtype=np.float64)
        w = C.c_double(float(om))
        liblb.bgk(g.ctypes.data_as(_P), C.cast(C.pointer(w), C.c_void_p))
        assert np.array_equal(g, want)


def test_boundary_operators_on_host_nodes(liblb):
    kat = np.load(os.path.join(GOLD, "kat.npz"))
    for which, name in enumerate(BC_NAMES):
        fn = getattr(liblb, name)
        for f, p, om, want in zip(kat["bc_in"][which], kat["bc_p"][which],
                                  kat["bc_omega"][which], kat["bc_out"][which]):
            w = C.c_double(float(om))
            bulk = Dynamics(C.cast(liblb.bgk, C.c_void_p).value, C.cast(C.pointer(w), C.c_void_p))
            if which == 0:
                data = None
            elif 5 <= which <= 8 or which >= 13:
                data = PressureBCData(C.pointer(bulk), float(p[0]), float(p[1]))
            else:
                data = VelocityBCData(C.pointer(bulk), float(p[0]), float(p[1]))
            g = np.array(f, dtype=np.float64)
            fn(g.ctypes.data_as(_P), None if data is None else C.cast(C.pointer(data), C.c_void_p))
            assert np.array_equal(g, want), name
