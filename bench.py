"""Benchmark of the D2Q9 BGK collide-stream hot path on one MI355X.

Default workload (N=1): the 16384x4096 cylinder channel (disc (lx/5, ly/2,
ly/10), Re 100, u_max 0.02, regularized walls, Poiseuille inlet, ZGB
pressure outlet), FP64 populations, W warm-up steps then K timed steps of the
fused pull kernel (the combine variant).  Prints ONE JSON line.

The populations are initialised on the device (equilibrium with the inlet
profile), so the timed region starts with everything resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "memory-aware-lbm_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "GLUPS & % of HBM roofline (FP64 vs FP32-store); serving p50/p99 latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--lx", type=int, default=16384)
    ap.add_argument("--ly", type=int, default=4096)
    ap.add_argument("--variant", default="combine")
    ap.add_argument("--variants", default="origin,combine",
                    help="variants timed besides the headline one (comma list, or '')")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "off"])
    return ap.parse_args()


def time_variant(lbm, ch, variant, steps, warmup):
    with lbm.Lattice(ch) as lat:
        lat.init_equilibrium()
        lat.step(warmup, variant)
        lat.sync()
        ms = lat.time_steps(steps, variant)
        bpn = lat.bytes_per_node(variant)
    cells = ch.lx * ch.ly
    glups = cells * steps / (ms * 1e-3) / 1e9
    return dict(ms=ms, ms_per_step=ms / steps, glups=glups, bytes_per_node=bpn,
                gbs=glups * bpn)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    import lbm_amd as lbm

    ch = lbm.Channel.geometry(args.lx, args.ly, "B")
    head = time_variant(lbm, ch, args.variant, args.steps, args.warmup)
    others = {}
    for v in [v for v in args.variants.split(",") if v and v != args.variant]:
        try:
            r = time_variant(lbm, ch, v, max(args.steps // 4, 10), min(args.warmup, 20))
            others[v] = {k: round(r[k], 4) for k in ("glups", "gbs", "ms_per_step",
                                                    "bytes_per_node")}
        except lbm.LbmError as e:
            others[v] = {"error": str(e)}
    achieved = head["gbs"]
    copy_gbs = lbm.copy_bandwidth(2 << 30, 20)
    out = {
        "metric": METRIC,
        "value": round(head["glups"], 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cylinder channel {args.lx}x{args.ly}, Re 100, FP64, "
                               f"{args.variant} (fused pull) kernel",
                   "lx": args.lx, "ly": args.ly, "variant": args.variant},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "variants": others,
        "copy_kernel_gbs": round(copy_gbs, 1),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
