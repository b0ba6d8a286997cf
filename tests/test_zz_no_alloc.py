"""lbm_hip_apply_node_op_noalloc (the node operators staged through a fixed
device array instead of a per-call allocation) against
lbm_hip_apply_node_op, through the C-ABI: bit-identical outputs for every
operator on two calls in a row, the second with fewer vectors (the array
keeps the first call's contents beyond them), the reference-written
known answers, and the size limit."""
import ctypes as C
import os

import numpy as np
import pytest

import lbm_amd as L

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
_P = C.POINTER(C.c_double)
MAX = 4096   # LBM_NODE_OP_MAX


def _fns():
    lib = L.library()
    fn = lib.lbm_hip_apply_node_op_noalloc
    fn.argtypes = [C.c_int, C.c_int, _P, _P, _P, C.c_double]
    fn.restype = C.c_int
    return lib.lbm_hip_apply_node_op, fn


def _call(fn, which, f, p0, p1, omega):
    g = np.ascontiguousarray(f, dtype=np.float64).copy()
    a = np.ascontiguousarray(p0, dtype=np.float64)
    b = np.ascontiguousarray(p1, dtype=np.float64)
    rc = fn(int(which), g.shape[0], g.ctypes.data_as(_P), a.ctypes.data_as(_P),
            b.ctypes.data_as(_P), float(omega))
    assert rc == 0, L.library().lbm_hip_last_error()
    return g


def _inputs(rng, which, n):
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    f = w * (1.0 + 0.05 * rng.standard_normal((n, 9)))
    if 5 <= which <= 8 or 13 <= which <= 16:      # pressure operators: (rho, uPar)
        p0 = 1.0 + 0.01 * rng.standard_normal(n)
    else:                                         # velocity operators: (ux, uy)
        p0 = 0.02 * rng.standard_normal(n)
    p1 = 0.01 * rng.standard_normal(n)
    return f, p0, p1


@pytest.mark.parametrize("which", range(18))
def test_noalloc_equals_allocating_entry(which):
    alloc, noalloc = _fns()
    rng = np.random.default_rng(100 + which)
    omega = 1.0 + rng.random()
    for n in (MAX, 777):                          # the second call smaller
        f, p0, p1 = _inputs(rng, which, n)
        want = _call(alloc, which, f, p0, p1, omega)
        got = _call(noalloc, which, f, p0, p1, omega)
        assert np.array_equal(got, want), (which, n)
        assert np.isfinite(got).all()


def test_noalloc_reference_known_answers():
    _, noalloc = _fns()
    kat = np.load(os.path.join(GOLD, "kat.npz"))
    for which in range(17):
        f, p, om = kat["bc_in"][which], kat["bc_p"][which], kat["bc_omega"][which]
        for k in range(len(f)):
            got = _call(noalloc, which, f[k:k + 1], p[k:k + 1, 0], p[k:k + 1, 1], om[k])
            assert np.array_equal(got[0], kat["bc_out"][which][k]), (which, k)
    f, om = kat["bgk_in"], kat["bgk_omega"]
    for k in range(len(f)):
        got = _call(noalloc, 17, f[k:k + 1], np.zeros(1), np.zeros(1), om[k])
        assert np.array_equal(got[0], kat["bgk_out"][k])


def test_noalloc_rejects_more_than_the_array_holds():
    _, noalloc = _fns()
    n = MAX + 1
    f = np.zeros((n, 9))
    z = np.zeros(n)
    rc = noalloc(17, n, f.ctypes.data_as(_P), z.ctypes.data_as(_P), z.ctypes.data_as(_P), 1.0)
    assert rc == -1
    assert "LBM_NODE_OP_MAX" in L.library().lbm_hip_last_error().decode()
