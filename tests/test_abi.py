"""The C-ABI library builds, loads and exports every function the public
headers declare (no device calls: these run without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "memory-aware-lbm_amd")
LIBS = {
    "lbm_hip.h": os.path.join(PKG, "liblbm_hip.so"),
    "lb.h": os.path.join(PKG, "liblb.so"),
    "boundaries.h": os.path.join(PKG, "liblb.so"),
}
# declared by lb.h but defined by the driver, as in the reference (unsteady.c:205)
DRIVER_DEFINED = {"updateZeroGradientBoundary"}


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        if name in ("if", "while", "for", "return", "sizeof", "void", "int", "double",
                    "char", "float", "long", "short", "unsigned", "const"):
            continue
        names.add(name)
    return names


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header", sorted(LIBS))
def test_header_symbols_exported(header):
    so = LIBS[header]
    assert os.path.exists(so), f"{so} not built"
    want = declared(header)
    assert want, header
    missing = want - exported(so) - DRIVER_DEFINED
    assert not missing, f"{header}: not exported by {os.path.basename(so)}: {sorted(missing)}"


def test_library_loads_without_gpu():
    import lbm_amd
    lib = lbm_amd.library()
    assert lib.lbm_hip_last_error() is not None


def test_footprint_without_allocating():
    import ctypes as C
    import lbm_amd
    lib = lbm_amd.library()
    p = lbm_amd._Problem()
    p.lx, p.ly, p.storage = 16384, 4096, lbm_amd.STORE_F64
    b = lib.lbm_hip_footprint_of(C.byref(p))
    # two FP64 population buffers of 9 planes each dominate
    assert 9.6e9 < b < 1.05e10


def test_footprint_of_a_sweep_counts_its_buffers():
    """Admission support: the four- and six-step sweeps allocate one and
    three further population buffers of 9 planes of (lx + 2) x pitch elements
    (the six-step sweep's eight-step launch, for remainders of 2 or 3 steps,
    takes the third)."""
    import lbm_amd
    for lx, ly, st, es in ((2048, 1024, "f64", 8), (300, 77, "f32", 4), (64, 5, "f16", 2)):
        ch = lbm_amd.Channel.geometry(lx, ly, "B")
        base = lbm_amd.footprint_of(ch, st)
        pitch = (63 + ly + 2 + 63) // 64 * 64
        buf = 9 * (lx + 2) * pitch * es
        assert lbm_amd.footprint_of(ch, st, "combine") == base
        assert lbm_amd.footprint_of(ch, st, "panel") == base
        assert lbm_amd.footprint_of(ch, st, "panel4") == base + buf
        assert lbm_amd.footprint_of(ch, st, "panel6") == base + 3 * buf


def test_unbound_sweep_aborts_without_gpu():
    """A sweep on a Simulation that was never bound to a device aborts with a
    message instead of computing anything on the CPU."""
    import sys
    code = (
        "import ctypes as C\n"
        f"lib = C.CDLL({os.path.join(PKG, 'liblb.so')!r})\n"
        "class Sim(C.Structure):\n"
        "    _fields_ = [('lx', C.c_int), ('ly', C.c_int), ('a', C.c_void_p), ('b', C.c_void_p),"
        " ('c', C.c_void_p), ('d', C.c_void_p)]\n"
        "s = Sim()\n"
        "lib.constructSim(C.byref(s), 8, 8)\n"
        "lib.collide_with_stream(C.byref(s))\n"
        "print('NOT REACHED')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode != 0
    assert "NOT REACHED" not in r.stdout
    assert "not bound to a GPU" in r.stderr


def test_driver_usage_message():
    exe = os.path.join(PKG, "unsteady_gpu")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def _valid_problem():
    import lbm_amd
    p = lbm_amd._Problem()
    p.lx, p.ly, p.obst_x, p.obst_y, p.obst_r = 64, 32, 12, 16, 3
    p.omega = 1.5
    for s in range(4):
        p.side[s].kind, p.side[s].family = lbm_amd.BC_VELOCITY, lbm_amd.BC_REGULARIZED
    p.side[lbm_amd.RIGHT].kind = lbm_amd.BC_PRESSURE
    p.zgb, p.storage = 1, lbm_amd.STORE_F64
    return p


@pytest.mark.parametrize("field,value,message", [
    ("lx", 2, "lx and ly must be >= 3"),
    ("ly", 0, "lx and ly must be >= 3"),
    ("lx", 65535, "lattice too large"),          # (lx+2)(ly+2) > 2^31 with ly = 32766
    ("omega", 0.0, "omega must lie in (0, 2]"),
    ("omega", 2.5, "omega must lie in (0, 2]"),
    ("storage", 7, "unknown storage format"),
    ("side_kind", 9, "side kind must be velocity or pressure"),
    ("zgb_left_only", 1, "zgb needs a pressure operator on the right side"),
])
def test_create_rejects_bad_problems_without_gpu(field, value, message):
    """lbm_hip_create validates the problem before any device call: the
    degenerate and oversized lattices, and the other invalid settings, come
    back as LBM_ERR_ARG with a message and no context."""
    import ctypes as C
    import lbm_amd
    lib = lbm_amd.library()
    p = _valid_problem()
    if field == "lx" and value == 65535:
        p.lx, p.ly = 65535, 32766
    elif field == "side_kind":
        p.side[0].kind = value
    elif field == "zgb_left_only":
        p.side[lbm_amd.RIGHT].kind = lbm_amd.BC_VELOCITY
    else:
        setattr(p, field, value)
    h = C.c_void_p(1234)
    rc = lib.lbm_hip_create(C.byref(p), C.byref(h))
    assert rc == -1                                        # LBM_ERR_ARG
    assert h.value is None
    assert message in lib.lbm_hip_last_error().decode()


def test_reference_executable_names_installed():
    """unsteady_<variant>[_openmp][_papi] for the reference's five sweeps
    (CMakeLists.txt:27-50) resolve to the GPU driver."""
    exe = os.path.realpath(os.path.join(PKG, "unsteady_gpu"))
    for v in ("origin", "combine", "tight", "tight_block", "panel"):
        for s in ("", "_openmp", "_papi", "_openmp_papi"):
            link = os.path.join(PKG, "bin", f"unsteady_{v}{s}")
            assert os.path.realpath(link) == exe, link
