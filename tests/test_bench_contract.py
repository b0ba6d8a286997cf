"""bench.py's own plumbing, on the CPU: the HBM traffic it reports is read
from the committed rocprofv3 summary, and runs on more than one GPU are
refused with a message, and the sweeps' timed regions run with launch events
(the headline) or without them, with a second pass for the launch times."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_panel_traffic_from_profile():
    b = load_bench()
    # interior + perimeter dispatches of one two-, four- or six-step launch on
    # 16384 x 4096: at least the algorithmic 144 B per node (one read and one
    # write of nine FP64 populations), at most 1.5x that
    alg = 144.0 * 16384 * 4096
    for v in ("panel", "panel4", "panel6"):
        t = b.measured_traffic(b.LAUNCH_KEY[(v, "f64")])
        assert t is not None and alg <= t <= 1.5 * alg, (v, t)
    # the interior march alone updates 98-99% of the nodes
    t = b.measured_traffic("k_wpanel_in<double, 2, 6>")
    assert t is not None and 0.95 * alg <= t <= 1.5 * alg, t
    assert b.measured_traffic("k_block<double, 64>") > alg


def test_cpu_info_names_the_host():
    b = load_bench()
    info = b.cpu_info()
    assert info["logical_cpus"] >= 1
    assert info["cpus_in_affinity"] >= 1
    assert info["physical_cores"] is None or info["physical_cores"] <= info["logical_cpus"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


class _FakeLattice:
    """Records the calls time_variant makes; each timed pass takes 2 ms with
    launch events and 1 ms without."""
    log = []

    def __init__(self, ch, storage="f64"):
        self.timing = False

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False

    def init_equilibrium(self):
        self.log.append("init")

    def step(self, n, variant):
        self.log.append(("step", n))

    def sync(self):
        pass

    def kernel_timing(self, on):
        self.timing = on
        self.log.append(("timing", on))

    def time_steps(self, n, variant):
        self.log.append(("time", n, self.timing))
        return 2.0 if self.timing else 1.0

    def kernel_times(self):
        return [{"kernel": "k", "depth": 4, "launches": 25, "avg_ms": 0.08, "min_ms": 0.07,
                 "max_ms": 0.09, "nodes": 100.0}]


class _FakeLbm:
    Lattice = _FakeLattice


class _Ch:
    lx, ly = 10, 10


def test_time_variant_probed_and_unprobed_passes():
    """The headline times its region with launch events; the other sweeps
    time theirs without and record launch times in a second pass of as many
    steps."""
    b = load_bench()
    _FakeLattice.log = []
    r = b.time_variant(_FakeLbm, _Ch, "panel4", 100, 5, probed=True)
    assert [e for e in _FakeLattice.log if e[0] == "time"] == [("time", 100, True)]
    assert r["ms"] == 2.0 and r["glups"] == r["glups_probed"]
    _FakeLattice.log = []
    r = b.time_variant(_FakeLbm, _Ch, "panel4", 100, 5, probed=False)
    assert [e for e in _FakeLattice.log if e[0] == "time"] == [("time", 100, False),
                                                               ("time", 100, True)]
    assert r["ms"] == 1.0 and r["glups"] == 2 * r["glups_probed"]
    assert r["kernel_avg_ms"] == 0.08 and r["steps_by_depth"] == {"4": 100}
