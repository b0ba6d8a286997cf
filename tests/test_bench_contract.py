"""bench.py's own plumbing, on the CPU: the HBM traffic it reports is read
from the committed rocprofv3 summary, and runs on more than one GPU are
refused with a message."""
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
