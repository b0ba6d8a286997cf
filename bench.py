"""Benchmark of the D2Q9 BGK collide-stream hot path on one MI355X.

Workload (N=1): the 16384x4096 cylinder channel -- disc (lx/5, ly/2, ly/10),
Re 100, u_max 0.02, regularized walls, Poiseuille inlet, zero-gradient
pressure outlet -- with FP64 populations, bit-identical to the reference C
code.  Every kernel's code is loaded first on a small channel; then, on the
full lattice, the headline sweep first runs `--settle` steps (123: about
70 ms, nineteen six-step launches and one eight-step launch, so the GPU's
clocks and translation caches are warm and the sweep's further buffers
allocated), the lattice is
re-initialised to the initial condition, then W untimed warm-up steps and K
timed steps (HIP events on the lattice's stream around exactly those K
steps); one JSON line on stdout.

`value` is the GLUPS of the headline sweep variant (default panel6: six time
steps per launch, five of them held in wave64 registers; the fastest FP64
sweep on this lattice): 16384 x 4096 x K / the timed region.  K steps of a
deep sweep are as many six-step launches as fit, then the launches of the
remainder (a remainder of 2 or 3 as one eight-step launch in place of the
last six-step one); `steps_by_depth` says how many steps ran at each depth.
The other variants are timed on the same lattice ("variants"), with the FP32
and FP16 storage forms ("storage_forms"): their `glups` over a timed region
without launch events, their kernel times (and `glups_with_launch_events`)
from a second pass of as many steps with them.  The populations are
initialised on the device, so everything is resident in HBM when the timed
region starts.

roofline: the dominant kernel is the one that carries the sweep's lattice
updates (for panel6 the interior march k_wpanel_in<double, 2, 6>, whose
perimeter kernels run beside it on a second stream).  Its average launch
duration is measured live with HIP events on its own stream over the timed
region.  achieved = the algorithmic bytes of one launch -- 144 B per node
(nine FP64 populations read once and written once; 72 B FP32, 36 B FP16
storage) times the nodes that kernel updates per launch -- divided by that
duration; frac = achieved / 8 TB/s.  A k-step launch performs k updates per
node for those bytes: the update-equivalent rate (achieved x k) is reported
beside it under its own name and is not a roofline figure.  traffic is the
same kernel's measured HBM bytes per launch, from rocprofv3 FETCH_SIZE and
WRITE_SIZE (profiles/kernels_16384x4096.md); traffic_launch adds the
perimeter and ring dispatches of the same launch.

cpu_baseline: the reference's own OpenMP sweeps (origin_openmp and
tight_block_openmp from its lb.c / boundaries.c, built unmodified into
oracle/_ref/libref_harness.so) on the same 16384x4096 disc-B channel, timed on
the host cores of this machine for a bounded sample (10 warm-up + 100 timed
steps by default, the GPU's 1:10 ratio), with the CPU model and core counts.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "memory-aware-lbm_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "GLUPS & % of HBM roofline (FP64 vs FP32-store); serving p50/p99 latency"
BYTES_PER_UPDATE = {"f64": 144.0, "f32": 72.0, "f16": 36.0}
PROFILE = os.path.join(ROOT, "profiles", "kernels_16384x4096.md")
# profile entry holding one whole launch (interior, perimeter and ring
# dispatches added up) of the panel sweeps
LAUNCH_KEY = {("panel", "f64"): "panel_launch<double>", ("panel4", "f64"): "panel4_launch<double>",
              ("panel6", "f64"): "panel6_launch<double>", ("panel", "f32"): "panel_launch<float>",
              ("panel4", "f32"): "panel4_launch<float>", ("panel", "f16"): "panel_launch<_Float16>",
              ("panel4", "f16"): "panel4_launch<_Float16>"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--lx", type=int, default=16384)
    ap.add_argument("--ly", type=int, default=4096)
    ap.add_argument("--variant", default="panel6")
    ap.add_argument("--variants", default="origin,combine,tight_block,panel,panel4",
                    help="variants timed besides the headline one (comma list, or '')")
    ap.add_argument("--other-steps", type=int, default=1000)
    ap.add_argument("--warm-code", type=int, default=1, choices=[0, 1],
                    help="load every kernel on a small channel first (0: not, so that a "
                         "profile holds only the full lattice's dispatches)")
    ap.add_argument("--settle", type=int, default=123,
                    help="steps the headline sweep runs on the lattice before it is "
                         "re-initialised for the warm-up and timed steps (0: none)")
    ap.add_argument("--storages", default="f32,f16",
                    help="reduced-precision storage forms, timed with the fused, the "
                         "tight_block, the panel and the four-step sweeps")
    ap.add_argument("--batch", type=int, default=512,
                    help="channels of the batching workload (0 = skip)")
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--batch-steps", type=int, default=96)
    ap.add_argument("--batch-tune", default="on", choices=["on", "off", "panel", "panel4", "panel6"],
                    help="autotune the batch launch shape (off: the library's defaults, "
                         "panel / panel4 / panel6: that register march at its default chunk "
                         "width)")
    ap.add_argument("--serve", type=int, default=40,
                    help="requests of the serving workload (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-seed", type=int, default=0)
    ap.add_argument("--serve-pool", type=int, default=1, choices=[0, 1],
                    help="contexts from a stream-ordered pool (0: hipMalloc / hipFree)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served window (no error re-runs) and print its "
                         "summary: for kernel traces of the serving loop")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "off"])
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="timed steps of each reference OpenMP sweep")
    ap.add_argument("--cpu-warmup", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads (0: OMP_NUM_THREADS if set, else every physical "
                         "core this process may run on)")
    return ap.parse_args()


def warm_code(lbm):
    """Load every sweep kernel's code object before any timed region: each
    variant once on a 700x400 channel, through every launch depth it uses
    (1 + 6 + 8 + 4 + 2 + 1 steps: the first collision, a six-, an eight-, a
    four- and a two-step launch, a single-step remainder), in every storage
    form."""
    ch = lbm.Channel.geometry(700, 400, "B")
    for storage in ("f64", "f32", "f16"):
        with lbm.Lattice(ch, storage=storage) as lat:
            for v in ("origin", "combine", "tight_block", "panel", "panel4", "panel6"):
                lat.init_equilibrium()
                for n in (7, 8, 4, 2, 1):
                    lat.step(n, v)
            lat.sync()


def dominant(recs):
    """The record of the kernel that takes most of the timed region (the
    six-step launches of a 20-step panel6 run, not its one eight-step
    launch); ties go to the deeper one."""
    return max(recs, key=lambda r: (r["launches"] * r["avg_ms"], r["depth"]))


def time_variant(lbm, ch, variant, steps, warmup, storage="f64", settle=0, probed=True):
    """Throughput of `steps` steps of a sweep and the launch times of its
    dominant kernel.  probed: the launch times come from HIP events recorded
    around every launch inside the timed region itself (the headline);
    otherwise the timed region runs without them and a second pass of as many
    steps records them.  The events change how a deep sweep's interior kernel
    overlaps its frame kernels on the other stream (FP32 four-step sweep on
    16384x4096: 149.6 GLUPS without them, 134.5 with), so the sweeps other
    than the headline report the throughput measured without them."""
    with lbm.Lattice(ch, storage=storage) as lat:
        if settle:
            # run the sweep on this lattice first so the GPU's clocks, its
            # translation caches and the sweep's further buffers are warm,
            # then start over from the initial condition
            lat.init_equilibrium()
            lat.step(settle, variant)
        lat.init_equilibrium()
        lat.step(max(warmup, 1), variant)   # leaves the state in the variant's convention
        lat.sync()
        if not probed:
            ms = lat.time_steps(steps, variant)
        lat.kernel_timing(True)
        ms_probed = lat.time_steps(steps, variant)
        recs = lat.kernel_times()
        lat.kernel_timing(False)
        if probed:
            ms = ms_probed
    cells = ch.lx * ch.ly
    dom = dominant(recs)
    bpu = BYTES_PER_UPDATE[storage]
    achieved = bpu * dom["nodes"] / (dom["avg_ms"] * 1e-3) / 1e9
    by_depth = {}
    for r in recs:
        by_depth[str(r["depth"])] = by_depth.get(str(r["depth"]), 0) + r["depth"] * r["launches"]
    return dict(ms=ms, ms_per_step=ms / steps, glups=cells * steps / (ms * 1e-3) / 1e9,
                glups_probed=cells * steps / (ms_probed * 1e-3) / 1e9,
                kernel=dom["kernel"], depth=dom["depth"], kernel_avg_ms=dom["avg_ms"],
                kernel_nodes=dom["nodes"], achieved=achieved, steps_by_depth=by_depth,
                launches=[{"kernel": r["kernel"], "depth": r["depth"], "launches": r["launches"],
                           "avg_ms": round(r["avg_ms"], 5), "min_ms": round(r["min_ms"], 5),
                           "max_ms": round(r["max_ms"], 5)} for r in recs])


def measured_traffic(key):
    """HBM bytes per launch of a kernel (or of a whole panel launch) from the
    committed rocprofv3 PMC summary, or None."""
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*(.*)", text)
    if not m:
        return None
    # names hold ", " themselves (template arguments): split at "=<digits>"
    table = {k.strip(): float(v) for k, v in re.findall(r"([^=]+)=(\d+)(?:,\s*)?", m.group(1))}
    return table.get(key)


def summary(r, storage):
    out = {"glups": round(r["glups"], 4), "ms_per_step": round(r["ms_per_step"], 5),
           "glups_with_launch_events": round(r["glups_probed"], 4),
           "kernel": r["kernel"], "steps_per_launch": r["depth"],
           "kernel_avg_ms": round(r["kernel_avg_ms"], 5),
           "nodes_per_launch": int(r["kernel_nodes"]),
           "achieved_gbs": round(r["achieved"], 1),
           "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
           "update_equivalent_gbs": round(r["achieved"] * r["depth"], 1),
           "steps_by_depth": r["steps_by_depth"]}
    return out


def batching(args):
    """The batching workload: seeded uneven channels, hipGraph replay,
    autotuned launch shape (two steps per launch over LDS tiles or by the
    register march over every channel's wave grid, four or six by the deep
    register marches, or one persistent launch per step); FP64, FP32 and
    FP16 storage."""
    from lbm_amd.batch import Batch, autotune, channels_from_seed
    chans = channels_from_seed(args.batch, seed=args.batch_seed)
    out = {"channels": len(chans), "seed": args.batch_seed,
           "cells": int(sum(c.lx * c.ly for c in chans)),
           "tau_below_0505": int(sum(1 for c in chans if 1.0 / c.omega() < 0.505))}
    for storage in ("f64", "f32", "f16"):
        with Batch(chans, storage) as b:
            b.init_equilibrium()
            if args.batch_tune == "on":
                tune = autotune(b, steps=40, graph_steps=10, use_cache=False)
            else:
                b.set_launch(0, 0, 12)
                if args.batch_tune != "off":
                    b.set_sweep(args.batch_tune)
                tune = dict(b.launch)
            spl = b.steps_per_launch()
            b.step(tune["graph_steps"] or 12)
            ms = b.time_steps(args.batch_steps)
            glups = b.cells * args.batch_steps / (ms * 1e-3) / 1e9
            bpu = BYTES_PER_UPDATE[storage]
            # a k-step launch reads and writes every population once (with a
            # step count that is a multiple of the graph's, every launch is k
            # steps deep)
            ach = glups * bpu / spl
            out[storage] = {"glups": round(glups, 4), "ms_per_step": round(ms / args.batch_steps, 5),
                            "achieved_gbs": round(ach, 1),
                            "frac": round(ach / HBM_PEAK_GBS, 4),
                            "update_equivalent_gbs": round(glups * bpu, 1),
                            "launch": {"depth": tune["depth"], "steps_per_launch": spl,
                                       "tile_cols": tune["tile_cols"],
                                       "ty_cap": tune["ty_cap"], "wgs_per_cu": tune["wgs_per_cu"],
                                       "graph_steps": tune["graph_steps"],
                                       "sweep": tune.get("sweep", "tight_block"),
                                       "chunk_cols": tune.get("chunk_cols", 0),
                                       "panel_cols": tune.get("panel_cols", 0)},
                            "footprint_gb": round(b.footprint() / 1e9, 3)}
    return out


def serving(args):
    """The serving workload: a seeded Poisson stream of solves of uneven
    size, length and storage, admitted by HBM footprint and run concurrently
    with device-resident lattices.  FP32 solves are re-run alone in FP64
    afterwards (outside the served window) for their velocity error."""
    import numpy as np
    import lbm_amd as lbm
    from lbm_amd.serve import SolveServer, poisson_requests, summarize
    reqs = poisson_requests(args.serve, args.serve_rate, seed=args.serve_seed)
    srv = SolveServer(max_active=8, slice_ms=20.0, graph_steps=50, keep_fields=True,
                      pool_bytes=(8 << 30) if args.serve_pool else None)
    res = srv.run(reqs)
    out = summarize(res)
    if args.serve_only:
        return out
    out = {k: round(v, 5) if isinstance(v, float) else v for k, v in out.items()}
    out.update({"rate_per_s": args.serve_rate, "seed": args.serve_seed,
                "max_active": srv.max_active, "hbm_budget_gb": round(srv.hbm_budget / 1e9, 1),
                "peak_active": srv.stats["peak_active"],
                "peak_resident_gb": round(srv.stats["peak_bytes"] / 1e9, 3),
                "f64_solves": sum(r.storage == "f64" for r in res),
                "f32_solves": sum(r.storage == "f32" for r in res)})
    errs = []
    for req, r in zip(reqs, res):
        if r.storage != "f32":
            continue
        with lbm.Lattice(req.channel, "f64") as lat:
            lat.init_equilibrium()
            lat.set_graph_steps(50)
            lat.step(req.steps, "combine")
            _, u, v = lat.macros()
        ref = np.stack([u, v])
        got = np.stack(r.fields[1:])
        errs.append(float(np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum())))
    if errs:
        out["f32_velocity_rel_l2_vs_f64"] = {"median": float(np.median(errs)),
                                             "max": float(np.max(errs))}
        # the worst FP32 solve again, FP32 and FP64 side by side: its velocity
        # error against the step count
        from lbm_amd.precision import lockstep
        f32 = [req for req in reqs if req.storage == "f32"]
        worst = f32[int(np.argmax(errs))]
        sample = max(1, worst.steps // 100)
        marks = [sample * 10 * k for k in range(1, 10)]
        r = lockstep(worst.channel, worst.steps, sample, ("f64", "f32"), checkpoints=marks)
        c = worst.channel
        out["worst_f32_solve"] = {
            "rid": worst.rid, "lattice": [c.lx, c.ly], "disc": [c.obst_x, c.obst_y, c.obst_r],
            "reynolds": round(c.reynolds, 3), "steps": worst.steps,
            "u_l2_rel_by_steps": {str(e["steps"]): e["u_l2_rel"] for e in r["errors"]["f32"]}}
    out["f64_velocity_error"] = "0 (bit-identical FP64 path)"
    return out


def cpu_info():
    """CPU model, physical cores and logical CPUs of the host, and how many of
    them this process may run on (its affinity mask)."""
    model, cores, logical = None, {}, 0
    proc = phys = core = None
    try:
        text = open("/proc/cpuinfo").read()
    except OSError:
        text = ""
    for block in text.strip().split("\n\n"):
        kv = {}
        for line in block.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        if "processor" not in kv:
            continue
        logical += 1
        model = model or kv.get("model name")
        proc, phys, core = int(kv["processor"]), kv.get("physical id", "0"), kv.get("core id", kv["processor"])
        cores[proc] = (phys, core)
    aff = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else set(cores)
    return {"model": model, "logical_cpus": logical or os.cpu_count(),
            "physical_cores": len(set(cores.values())) or None,
            "cpus_in_affinity": len(aff),
            "physical_cores_in_affinity": len({cores[p] for p in aff if p in cores}) or None}


def cpu_baseline(args):
    """The reference's OpenMP sweeps on the host cores, bounded sample."""
    so = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")
    if not os.path.exists(so):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    info = cpu_info()
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) \
        or info["physical_cores_in_affinity"] or 1
    # OpenMP reads these when the runtime loads with the harness below
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    lib = C.CDLL(so)
    lib.ref_time.argtypes = [C.POINTER(ol._Cfg), C.c_int, C.POINTER(C.c_double)]
    lib.ref_time.restype = C.c_int
    ch = ol.Channel.geometry(args.lx, args.ly, "B")
    tb = max(64, (args.lx // threads) // 64 * 64)   # thread_block: a multiple of blk_size
    res = {}
    for name, variant in (("origin_openmp", 0), ("tight_block_openmp", 3)):
        cfg = ol._Cfg(ch.lx, ch.ly, ch.obst_x, ch.obst_y, ch.obst_r, ch.u_max, ch.reynolds,
                      0, variant, 1, threads, 64, tb, args.cpu_steps, 1)
        secs = C.c_double()
        ran = lib.ref_time(C.byref(cfg), args.cpu_warmup, C.byref(secs))
        if ran <= 0 or secs.value <= 0:
            continue
        res[name] = dict(glups=args.lx * args.ly * ran / secs.value / 1e9, t_total_s=secs.value,
                         steps=ran)
    if not res:
        return None
    best = max(res, key=lambda k: res[k]["glups"])
    return {"value": round(res[best]["glups"], 5), "unit": "GLUPS", "cores": threads,
            "kind": "reference",
            "sample": f"reference {best} sweep (lb.c, boundaries.c built unmodified), "
                      f"{args.lx}x{args.ly}, disc (lx/5, ly/2, ly/10), Re 100, "
                      f"{args.cpu_warmup} warm-up + {res[best]['steps']} timed steps, "
                      f"blk 64, thread_block {tb}, {threads} OpenMP threads "
                      f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                      f"OMP_PLACES={os.environ.get('OMP_PLACES')})",
            "cpu": info,
            "drivers": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in res.items()}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(how the path would shard is a design note only)"}))
        sys.exit(2)
    import lbm_amd as lbm

    if args.serve_only:
        print(json.dumps({"serving": serving(args)}), flush=True)
        return
    if args.warm_code:
        warm_code(lbm)
    ch = lbm.Channel.geometry(args.lx, args.ly, "B")
    head = time_variant(lbm, ch, args.variant, args.steps, args.warmup, settle=args.settle)
    others = {}
    for v in [v for v in args.variants.split(",") if v and v != args.variant]:
        others[v] = summary(time_variant(lbm, ch, v, args.other_steps, min(args.warmup, 20),
                                         probed=False), "f64")
    storage = {}
    for sname in [v for v in args.storages.split(",") if v]:
        storage[sname] = {"bytes_per_update": BYTES_PER_UPDATE[sname]}
        for v in ("combine", "tight_block", "panel", "panel4"):
            r = time_variant(lbm, ch, v, args.other_steps, min(args.warmup, 20), sname,
                             probed=False)
            storage[sname][v] = summary(r, sname)
            f64 = head if v == args.variant else None
            base = (f64 or {}).get("glups") or others.get(v, {}).get("glups")
            if base:
                storage[sname][v]["vs_f64_same_sweep"] = round(r["glups"] / base, 3)
    batch = batching(args) if args.batch > 0 else None
    serve = serving(args) if args.serve > 0 else None
    copy_gbs = lbm.copy_bandwidth(2 << 30, 20)
    full = (args.lx, args.ly) == (16384, 4096)
    traffic = measured_traffic(head["kernel"]) if full else None
    launch_key = LAUNCH_KEY.get((args.variant, "f64"))
    traffic_launch = measured_traffic(launch_key) if full and launch_key else None
    avg_s = head["kernel_avg_ms"] * 1e-3
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
        "config": {"workload": f"cylinder channel {args.lx}x{args.ly}, Re 100, FP64 "
                               f"populations, {args.variant} sweep",
                   "lx": args.lx, "ly": args.ly, "variant": args.variant,
                   "disc": [ch.obst_x, ch.obst_y, ch.obst_r]},
        "steps_by_depth": head["steps_by_depth"],
        "launches": head["launches"],
        "roofline": {"bound": "hbm", "achieved": round(head["achieved"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(head["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": head["kernel"],
                     "kernel_avg_ms": round(head["kernel_avg_ms"], 5),
                     "kernel_launches": next(r["launches"] for r in head["launches"]
                                             if r["kernel"] == head["kernel"]
                                             and r["depth"] == head["depth"]),
                     "nodes_per_launch": int(head["kernel_nodes"]),
                     "algorithmic_bytes_per_launch": int(144 * head["kernel_nodes"]),
                     "steps_per_launch": head["depth"],
                     "update_equivalent_gbs": round(head["achieved"] * head["depth"], 1),
                     "hbm_frac_measured": (round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4)
                                           if traffic else None),
                     "traffic_launch": traffic_launch},
        "variants": others,
        "storage_forms": storage,
        "batching": batch,
        "serving": serve,
        "copy_kernel_gbs": round(copy_gbs, 1),
        "cpu_baseline": cpu_baseline(args) if args.cpu_baseline == "auto" else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
