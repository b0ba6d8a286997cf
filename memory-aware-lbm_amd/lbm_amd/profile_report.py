"""Per-kernel time and HBM traffic report from rocprofv3 output.

Inputs are three rocprofv3 runs of the same command (CSV output):

    rocprofv3 --kernel-trace --stats --output-format csv -d <trace> -o run -- <cmd>
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch> -o run -- <cmd>
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <write> -o run -- <cmd>

and, optionally, the JSON line bench.py printed in the traced run (with
--batch-cells N: only the batch kernels of a batching trace of N cells and
whole launch times, from the trace alone).  The report
gives, per kernel, the dispatch count and the average / minimum duration, the
algorithmic bytes of one dispatch (bytes per lattice update -- 144 FP64, 72
FP32, 36 FP16 -- times the nodes the dispatch updates once, i.e. one read and
one write of every population) where the kernel carries a sweep's updates,
and the measured HBM bytes per dispatch: FETCH_SIZE and WRITE_SIZE (KiB)
scaled by what each reads for the plain copy kernel k_copy, whose bytes are
known.  The panel sweeps launch their perimeter kernels (k_wpanel_edge; the
deep sweeps' frame passes, k_frame) on a second stream before the interior
kernel (k_wpanel_in); each such dispatch is attributed to the next interior
dispatch, and one whole launch's traffic is reported as
panel_launch<S> / panel4_launch<S> / panel6_launch<S>.  The last line,
`traffic_bytes_per_launch: name=bytes, ...`, is what bench.py reads.

    python -m lbm_amd.profile_report --trace D1 --fetch D2 --write D3 \\
        [--bench bench.json] [--nodes 67108864] [--command "..."] > report.md
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
from collections import OrderedDict, defaultdict
from typing import Dict, List, Tuple

COPY_BYTES = 2 << 30        # k_copy moves 2 GiB in and 2 GiB out per dispatch (bench.py)
BPU = {"double": 144.0, "float": 72.0, "_Float16": 36.0}


_TYPES = {"d": "double", "f": "float", "DF16_": "_Float16", "i": "int", "j": "unsigned int"}


def demangle(sym: str) -> str:
    """The mangled names rocprofv3 leaves as they are (templates over
    _Float16, 'DF16_'): '_ZN3lbm<n><name>I<args>E...' -> 'name<args>', for
    the argument kinds this library's kernels use (types, integer and bool
    literals); anything else is returned unchanged."""
    m = re.match(r"_ZN3lbm(\d+)", sym)
    if not m:
        return sym
    n = int(m.group(1))
    k = m.end()
    name = sym[k:k + n]
    k += n
    if k >= len(sym) or sym[k] != "I":
        return name
    k += 1
    args = []
    while k < len(sym) and sym[k] != "E":
        lit = re.match(r"L([a-z])(\d+)E", sym[k:])
        if lit:
            kind, val = lit.group(1), lit.group(2)
            args.append(("true" if val == "1" else "false") if kind == "b" else val)
            k += lit.end()
            continue
        for code, t in _TYPES.items():
            if sym.startswith(code, k):
                args.append(t)
                k += len(code)
                break
        else:
            return sym
    return f"{name}<{', '.join(args)}>"


def short(name: str) -> str:
    """'void lbm::k_wpanel_in<double, 2, 6>(lbm::Problem, ...)' -> 'k_wpanel_in<double, 2, 6>'"""
    n = name.strip()
    if n.startswith("_Z"):
        return demangle(n)
    if n.startswith("void "):
        n = n[5:]
    n = n.replace("lbm::", "")
    depth, out = 0, []
    for ch in n:
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    return "".join(out).strip()


def _rows(d: str, suffix: str) -> List[dict]:
    rows = []
    for path in sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)):
        with open(path) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def dispatches(trace_dir: str) -> List[Tuple[int, str, float]]:
    """(dispatch id, kernel, duration ms) in submission order."""
    out = []
    for r in _rows(trace_dir, "kernel_trace.csv"):
        ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
        out.append((int(r["Dispatch_Id"]), short(r["Kernel_Name"]), ms))
    return sorted(out)


def counters(pmc_dir: str, counter: str) -> List[Tuple[int, str, float]]:
    """(dispatch id, kernel, counter value) in submission order."""
    acc: Dict[int, list] = {}
    for r in _rows(pmc_dir, "counter_collection.csv"):
        if not r["Counter_Name"].startswith(counter):
            continue
        k = int(r["Dispatch_Id"])
        if k not in acc:
            acc[k] = [short(r["Kernel_Name"]), 0.0]
        acc[k][1] += float(r["Counter_Value"])
    return sorted((k, v[0], v[1]) for k, v in acc.items())


def launch_key(interior: str) -> str:
    m = re.match(r"k_wpanel_in<([^,]+), \d+, (\d+)>", interior)
    if not m:
        return ""
    depth = {"2": "panel", "4": "panel4", "6": "panel6", "8": "panel8"}[m.group(2)]
    return f"{depth}_launch<{m.group(1)}>"


def groups(disp: List[Tuple[int, str, float]]) -> Dict[str, List[float]]:
    """Per launch key, the summed value of every whole launch: each perimeter
    dispatch (the two-step sweep's k_wpanel_edge, the deep sweeps' frame
    passes k_frame) belongs to the next interior dispatch."""
    out: Dict[str, List[float]] = defaultdict(list)
    pending = 0.0
    for _, name, v in disp:
        if name.startswith(("k_wpanel_edge", "k_frame")):
            pending += v
        elif name.startswith("k_wpanel_in"):
            out[launch_key(name)].append(pending + v)
            pending = 0.0
        else:
            pending = 0.0
    return out


def idle_gaps(trace_dir: str) -> Tuple[float, float, List[Tuple[float, float]]]:
    """Device-wide idle time of a kernel trace: the gaps between the union of
    all kernels' [start, end) intervals, on every stream together.  Returns
    (span ms, busy ms, [(gap start ms from the first kernel, gap ms), ...])."""
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                for r in _rows(trace_dir, "kernel_trace.csv"))
    if not iv:
        return 0.0, 0.0, []
    t0, gaps, busy = iv[0][0], [], 0
    cs, ce = iv[0]
    for s, e in iv[1:]:
        if s > ce:
            busy += ce - cs
            gaps.append(((ce - t0) * 1e-6, (s - ce) * 1e-6))
            cs, ce = s, e
        else:
            ce = max(ce, e)
    busy += ce - cs
    return (ce - t0) * 1e-6, busy * 1e-6, gaps


def gap_report(trace_dir: str, label: str) -> List[str]:
    span, busy, gaps = idle_gaps(trace_dir)
    big = sorted(gaps, key=lambda g: -g[1])
    out = [f"### {label}", "",
           f"{span:.1f} ms from the first kernel's start to the last one's end; some kernel "
           f"running for {busy:.1f} ms ({100 * busy / span if span else 0:.1f}%).  Idle gaps: "
           f"{sum(1 for g in gaps if g[1] >= 0.1)} of at least 0.1 ms, "
           f"{sum(1 for g in gaps if g[1] >= 1.0)} of at least 1 ms, "
           f"{sum(1 for g in gaps if g[1] >= 10.0)} of at least 10 ms; idle in gaps of at "
           f"least 1 ms: {sum(g[1] for g in gaps if g[1] >= 1.0):.1f} ms.", "",
           "| longest gaps: at ms | gap ms |", "|---|---|"]
    out += [f"| {t:.1f} | {g:.3f} |" for t, g in big[:10]]
    return out + [""]


def march_levels(name: str) -> int:
    """Steps of a k_batch_win<S, DEPTH, LEVELS> dispatch (2 without the third
    template argument)."""
    args = name[name.find("<") + 1:name.rfind(">")].split(",")
    return int(args[2]) if len(args) >= 3 else 2


def batch_marches(trace_dir: str) -> Dict[Tuple[str, int], List[float]]:
    """Whole launches of the batched register marches, per (storage type,
    steps).  A launch's dispatches (a two-step launch: k_batch_wedge on the
    second stream beside k_batch_win<S, 2, 2>; a four- or six-step launch:
    k_batch_win<S, 2, 4 or 6> beside its frame passes, k_batch_frame, and the
    two-step passes of its small channels) overlap in time, and the next
    launch starts only after all of them have ended (it forks from the join
    of the last one), so a launch is a maximal run of time-overlapping
    dispatches; it lasts from the earliest start to the latest end (ms) and
    has as many steps as its interior dispatch with the most levels."""
    rows = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"]))
                  for r in _rows(trace_dir, "kernel_trace.csv"))
    rows = [r for r in rows if r[2].startswith(("k_batch_wedge", "k_batch_win", "k_batch_frame"))]
    out: Dict[Tuple[str, int], List[float]] = defaultdict(list)

    def close(group):
        wins = [n for _, _, n in group if n.startswith("k_batch_win")]
        if wins:
            lv = max(march_levels(n) for n in wins)
            s, e = min(g[0] for g in group), max(g[1] for g in group)
            out[(stype(wins[0]), lv)].append((e - s) * 1e-6)

    group: List[Tuple[int, int, str]] = []
    end = 0
    for t0, t1, name in rows:
        if group and t0 >= end:
            close(group)
            group = []
        group.append((t0, t1, name))
        end = max(end, t1) if len(group) > 1 else t1
    if group:
        close(group)
    return out


def batch_report(trace_dir: str, cells: int) -> List[str]:
    kt = per_kernel(dispatches(trace_dir))
    out = ["## Kernel durations (rocprofv3 --kernel-trace)", "",
           "| kernel | calls | avg ms | min ms |", "|---|---|---|---|"]
    for name, v in sorted(kt.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| {name} | {len(v)} | {sum(v) / len(v):.4f} | {min(v):.4f} |")
    out += ["", "## Whole launches", "",
            f"k updates of all {cells:,} cells per k-step launch; algorithmic bytes per launch = "
            "bytes per lattice update (FP64 144, FP32 72, FP16 36) x cells (one read and one "
            "write of every population).  A register march launch lasts from the first start to "
            "the last end of its dispatches; an LDS-tile launch is one k_batch_block dispatch "
            "(two steps).", "",
            "| launch | steps | launches | avg ms | min ms | algorithmic GB/launch | "
            "algorithmic GB/s | GLUPS |", "|---|---|---|---|---|---|---|---|"]
    rows = [(f"march<{t}>", lv, v) for (t, lv), v in sorted(batch_marches(trace_dir).items())]
    rows += [(n, 2, v) for n, v in kt.items() if n.startswith("k_batch_block")]
    for name, lv, v in rows:
        avg = sum(v) / len(v)
        gb = BPU[stype(name)] * cells / 1e9
        out.append(f"| {name} | {lv} | {len(v)} | {avg:.4f} | {min(v):.4f} | {gb:.3f} | "
                   f"{gb / (avg * 1e-3):.1f} | {lv * cells / (avg * 1e-3) / 1e9:.2f} |")
    return out


def per_kernel(disp) -> "OrderedDict[str, List[float]]":
    out: "OrderedDict[str, List[float]]" = OrderedDict()
    for _, name, v in disp:
        out.setdefault(name, []).append(v)
    return out


def updates_nodes(name: str, nodes: int, interior_nodes: Dict[str, int]) -> int:
    """Nodes one dispatch updates once (0: not a sweep kernel)."""
    if name.startswith("k_wpanel_in"):
        return interior_nodes.get(name, 0)
    if name.startswith(("k_collide", "k_block", "k_stream")):
        return nodes
    return 0


def stype(name: str) -> str:
    for t in ("_Float16", "float", "double"):
        if f"<{t}" in name:
            return t
    return "double"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--gaps", action="append", default=[], metavar="LABEL=DIR",
                    help="instead: device-wide idle gaps of kernel traces (repeatable)")
    ap.add_argument("--batch-cells", type=int, default=0,
                    help="instead: the batch kernels of a batching trace of that many cells")
    ap.add_argument("--bench", default=None, help="bench.py's JSON line of the traced run")
    ap.add_argument("--nodes", type=int, default=16384 * 4096)
    ap.add_argument("--title", default="Per-kernel time and HBM traffic, 16384x4096 cylinder "
                                        "channel (FP64, FP32, FP16 storage)")
    ap.add_argument("--command", action="append", default=[],
                    help="a command that produced the inputs (repeatable)")
    a = ap.parse_args(argv)
    if a.gaps:
        out = [f"# {a.title}", ""]
        if a.command:
            out += ["Commands (from the repository root on one MI355X):", "", "```",
                    "cd /tmp && export TMPDIR=/tmp && cd /root/repo"] + a.command + ["```", ""]
        for spec in a.gaps:
            label, d = spec.split("=", 1)
            out += gap_report(d, label)
        print("\n".join(out))
        return

    if a.batch_cells:
        out = [f"# {a.title}", ""]
        if a.command:
            out += ["Commands (from the repository root on one MI355X):", "", "```",
                    "cd /tmp && export TMPDIR=/tmp && cd /root/repo"] + a.command + ["```", ""]
        out += batch_report(a.trace, a.batch_cells)
        if a.bench:
            text = open(a.bench).read().strip().splitlines()
            bl = json.loads(text[-1]).get("batching", {}) if text else {}
            out += ["", "The traced run's own bench line (HIP events over the timed graph "
                    "replays): " + ", ".join(f"{st} {bl[st]['glups']} GLUPS"
                                             for st in ("f64", "f32", "f16") if st in bl) + "."]
        print("\n".join(out))
        return

    bench = None
    if a.bench:
        text = open(a.bench).read().strip().splitlines()
        bench = json.loads(text[-1]) if text else None
    # nodes the interior kernels update per launch, from the bench line
    interior_nodes = {}
    if bench:
        rl = bench.get("roofline", {})
        if rl.get("kernel"):
            interior_nodes[rl["kernel"]] = int(rl["nodes_per_launch"])
        forms = [bench.get("variants") or {}] + list((bench.get("storage_forms") or {}).values())
        for form in forms:
            for v in form.values():
                if isinstance(v, dict) and "kernel" in v and "nodes_per_launch" in v:
                    interior_nodes[v["kernel"]] = int(v["nodes_per_launch"])
    disp = dispatches(a.trace)
    fetch = counters(a.fetch, "FETCH_SIZE")
    write = counters(a.write, "WRITE_SIZE")
    kt = per_kernel(disp)
    kf, kw = per_kernel(fetch), per_kernel(write)
    cf = COPY_BYTES / (sum(kf["k_copy"]) / len(kf["k_copy"]) * 1024) if "k_copy" in kf else 1.0
    cw = COPY_BYTES / (sum(kw["k_copy"]) / len(kw["k_copy"]) * 1024) if "k_copy" in kw else 1.0

    out = [f"# {a.title}", ""]
    if a.command:
        out += ["Commands (from the repository root on one MI355X):", "", "```",
                "cd /tmp && export TMPDIR=/tmp && cd /root/repo"] + a.command + ["```", ""]
    out += ["## Kernel durations (rocprofv3 --kernel-trace)", "",
            "Algorithmic bytes per dispatch = bytes per lattice update (FP64 144, FP32 72, "
            "FP16 36) x the nodes one dispatch updates (one read and one write of every "
            "population; a k-step launch updates them k times for those bytes).", "",
            "| kernel | calls | avg ms | min ms | algorithmic GB/dispatch | algorithmic GB/s "
            "| fraction of 8 TB/s |", "|---|---|---|---|---|---|---|"]
    for name, v in sorted(kt.items(), key=lambda kv: -sum(kv[1])):
        avg, mn = sum(v) / len(v), min(v)
        n = updates_nodes(name, a.nodes, interior_nodes)
        if n:
            gb = BPU[stype(name)] * n / 1e9
            out.append(f"| {name} | {len(v)} | {avg:.4f} | {mn:.4f} | {gb:.3f} | "
                       f"{gb / (avg * 1e-3):.1f} | {gb / (avg * 1e-3) / 8000:.3f} |")
        else:
            out.append(f"| {name} | {len(v)} | {avg:.4f} | {mn:.4f} |  |  |  |")
    out += ["", "## HBM traffic per dispatch (rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE)",
            "", f"Calibration on the plain copy kernel k_copy (2 GiB read + 2 GiB written per "
            f"dispatch): FETCH_SIZE scale x{cf:.4f}, WRITE_SIZE scale x{cw:.4f}.", "",
            "| kernel | FETCH KiB (raw) | WRITE KiB (raw) | HBM read GB | HBM write GB | "
            "traffic GB/dispatch |", "|---|---|---|---|---|---|"]
    traffic: "OrderedDict[str, float]" = OrderedDict()
    for name in kt:
        if name not in kf or name not in kw:
            continue
        f = sum(kf[name]) / len(kf[name])
        w = sum(kw[name]) / len(kw[name])
        rd, wr = f * 1024 * cf, w * 1024 * cw
        traffic[name] = rd + wr
        out.append(f"| {name} | {f:.0f} | {w:.0f} | {rd / 1e9:.3f} | {wr / 1e9:.3f} | "
                   f"{(rd + wr) / 1e9:.3f} |")
    gf, gw = groups([(k, n, v * 1024 * cf) for k, n, v in fetch]), \
        groups([(k, n, v * 1024 * cw) for k, n, v in write])
    out += ["", "Whole panel launches (the interior dispatch and the perimeter and ring "
            "dispatches launched for it):", "", "| launch | launches | HBM read GB | "
            "HBM write GB | traffic GB/launch |", "|---|---|---|---|---|"]
    for key in sorted(set(gf) & set(gw)):
        if not key:
            continue
        rd = sum(gf[key]) / len(gf[key])
        wr = sum(gw[key]) / len(gw[key])
        traffic[key] = rd + wr
        out.append(f"| {key} | {len(gf[key])} | {rd / 1e9:.3f} | {wr / 1e9:.3f} | "
                   f"{(rd + wr) / 1e9:.3f} |")
    if bench:
        rl = bench["roofline"]
        t = kt.get(rl["kernel"], [])
        out += ["", "## The traced run's own bench line", "",
                f"{bench['value']} GLUPS over {bench['steps']} timed steps "
                f"(steps by launch depth {bench.get('steps_by_depth')}); dominant kernel "
                f"{rl['kernel']}: {rl['kernel_avg_ms']} ms per launch measured with HIP events "
                f"over the timed region ({rl['kernel_launches']} launches), "
                f"{sum(t) / len(t) if t else float('nan'):.4f} ms average over all "
                f"{len(t)} of its traced dispatches (warm-up included); achieved "
                f"{rl['achieved']} GB/s, frac {rl['frac']}."]
    out += ["", "traffic_bytes_per_launch: " +
            ", ".join(f"{k}={int(round(v))}" for k, v in traffic.items())]
    print("\n".join(out))


if __name__ == "__main__":
    main()
