"""Memory-hierarchy counter report of a sweep, in the reference's format.

The reference builds with PAPI and prints, after its time loop, four event
totals and their per-thread averages (unsteady.c:533-548, events lb.h:43-44:
L1, L2 and L3 total cache misses and data-TLB misses):

    Global_E[00]: <total>
    ...
    AE: <total/threads> ...

and its scripts grep the `AE:` line (results/log.sh).  This tool prints the
same lines for a sweep on the GPU, from rocprofv3 counters summed over the
sweep's kernel dispatches, with the nearest counterpart of each event:

    E0  L1 (TCP) misses     TCP_TCC_READ_REQ_sum + TCP_TCC_WRITE_REQ_sum
                            (requests the vector L1 passes on to L2)
    E1  L2 (TCC) misses     TCC_MISS_sum
    E2  last level -> DRAM  TCC_EA0_RDREQ_DRAM_sum + TCC_EA0_WRREQ_DRAM_sum
                            (requests from L2 to HBM)
    E3  TLB misses          TCP_UTCL1_TRANSLATION_MISS_sum

`AE` divides by the number of compute units (the reference divides by its
OpenMP threads).  The time line is the reference's too
(`time eclapsed in N steps, t_total=...`, unsteady.c:558-559), measured
without the profiler.

    python -m lbm_amd.counters --variant panel4 --lx 16384 --ly 4096 --steps 20

runs the sweep three times: once for the time, then once per counter group
under `rocprofv3 --pmc` (the groups stay within one pass's hardware limits).
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import subprocess
import sys
import tempfile
from typing import Dict, List

EVENTS = [
    ("TCP_TCC_READ_REQ_sum", "TCP_TCC_WRITE_REQ_sum"),
    ("TCC_MISS_sum",),
    ("TCC_EA0_RDREQ_DRAM_sum", "TCC_EA0_WRREQ_DRAM_sum"),
    ("TCP_UTCL1_TRANSLATION_MISS_sum",),
]
# one rocprofv3 pass each (at most 4 TCP and 4 TCC hardware counters a pass)
PASSES = [
    ["TCP_TCC_READ_REQ_sum", "TCP_TCC_WRITE_REQ_sum", "TCP_UTCL1_TRANSLATION_MISS_sum"],
    ["TCC_MISS_sum"],
    ["TCC_EA0_RDREQ_DRAM_sum", "TCC_EA0_WRREQ_DRAM_sum"],
]
# kernels that belong to a sweep (set-up kernels such as k_init are left out)
SWEEP_KERNELS = ("k_collide", "k_stream", "k_block", "k_wpanel", "k_bump")


def _args(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--variant", default="panel4")
    ap.add_argument("--storage", default="f64")
    ap.add_argument("--lx", type=int, default=16384)
    ap.add_argument("--ly", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--rocprof", default="rocprofv3")
    ap.add_argument("--pass-timeout", type=int, default=120,
                    help="seconds before a counter pass is killed")
    return ap.parse_args(argv)


def run_sweep(a, warmup: int = 0) -> float:
    """Initialise the channel, run `warmup` untimed steps, then the sweep;
    return the sweep's seconds (HIP events on the lattice's stream)."""
    from . import Channel, Lattice
    ch = Channel.geometry(a.lx, a.ly, "B")
    with Lattice(ch, storage=a.storage) as lat:
        lat.init_equilibrium()
        if warmup:
            lat.step(warmup, a.variant)
        return lat.time_steps(a.steps, a.variant) / 1e3


def sum_counters(csv_path: str) -> Dict[str, float]:
    """Counter totals over the sweep kernels' dispatches of one pass."""
    out: Dict[str, float] = {}
    with open(csv_path) as fh:
        for r in csv.DictReader(fh):
            if not any(k in r["Kernel_Name"] for k in SWEEP_KERNELS):
                continue
            out[r["Counter_Name"]] = out.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def report(totals: Dict[str, float], units: int, steps: int, seconds: float) -> List[str]:
    """The reference's lines: Global_E per event, AE averages, time line."""
    ev = [sum(totals.get(c, 0.0) for c in names) for names in EVENTS]
    lines = [f"Global_E[{i:02d}]: {int(v)}" for i, v in enumerate(ev)]
    lines.append("AE: " + " ".join(str(int(v // units)) for v in ev) + " ")
    lines.append(f"time eclapsed in {steps} steps, t_total={seconds:.3f}, t_collid=0.000, "
                 f"t_stream=0.000, t_zgb=0.000")
    return lines


def _compute_units() -> int:
    try:
        import torch  # counting devices does not initialise the GPU here
        if torch.cuda.is_available():
            return int(torch.cuda.get_device_properties(0).multi_processor_count)
    except Exception:
        pass
    return 256


def main(argv=None):
    a = _args(argv)
    if a.child:
        run_sweep(a)
        return
    me = [sys.executable, "-m", "lbm_amd.counters", "--child", "--variant", a.variant,
          "--storage", a.storage, "--lx", str(a.lx), "--ly", str(a.ly), "--steps", str(a.steps)]
    env = dict(os.environ)
    pkg = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env["PYTHONPATH"] = pkg + os.pathsep + env.get("PYTHONPATH", "")
    # time without the profiler, in a child like the counter passes
    timer = [sys.executable, "-c",
             "import sys, lbm_amd.counters as c; a = c._args(sys.argv[1:]); "
             "print(c.run_sweep(a, warmup=4))", "--variant", a.variant, "--storage", a.storage,
             "--lx", str(a.lx), "--ly", str(a.ly), "--steps", str(a.steps)]
    seconds = float(subprocess.run(timer, env=env, check=True, capture_output=True,
                                   text=True).stdout.strip().splitlines()[-1])
    totals: Dict[str, float] = {}
    with tempfile.TemporaryDirectory() as tmp:
        for k, counters in enumerate(PASSES):
            d = os.path.join(tmp, f"pass{k}")
            cmd = ["timeout", "-s", "KILL", str(a.pass_timeout), a.rocprof, "--pmc", *counters,
                   "--output-format", "csv", "-d", d, "-o", "run", "--"] + me
            subprocess.run(cmd, env=env, check=True, capture_output=True)
            for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
                for name, v in sum_counters(path).items():
                    totals[name] = totals.get(name, 0.0) + v
    print(f"# {a.variant} sweep, {a.storage}, {a.lx}x{a.ly}, {a.steps} steps")
    for line in report(totals, _compute_units(), a.steps, seconds):
        print(line)


if __name__ == "__main__":
    main()
