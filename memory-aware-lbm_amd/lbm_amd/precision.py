"""Storage-precision study: FP32 and FP16 deviation storage against the
bit-exact FP64 path on the same channel.

Workload: the 16384 x 4096 cylinder channel, disc (lx/5, ly/2, ly/10), at
Re 100 and Re 1000, 20000 steps each.  The three storage forms start from
the same initial state and advance in lockstep, each on its fastest sweep
(FP64 six steps per launch, FP32 and FP16 four; every sweep gives the same
bits within a format, so the choice changes no number here).  Drag and lift
(momentum exchange on the disc, FP64 reduction) are sampled every `sample`
steps; the velocity fields are compared at a few checkpoints, so the error's
growth with the step count is recorded, and at the end.

Reported per (Re, storage):
  * velocity: L2 relative error ||u - u64|| / ||u64|| over both components,
    and max relative error max|u - u64| / max|u64|, at the end and at each
    checkpoint;
  * drag / lift series: max |d - d64| / max |d64| (same for lift), with the
    series' own scale (max |d64|, max |l64|);
  * mass: max |m - m64| / max |m64| over the series, and each run's own
    drift relative to its start (the channel is open, so FP64 drifts too).

The FP64 run is the reference of the comparison; it equals the reference C
code bit for bit (tests/test_gpu_parity.py, tests/test_gpu_bench_lattice.py),
and a CPU run of this size and length (1.3e12 lattice updates) is out of
reach.

    python -m lbm_amd.precision [--lx 16384 --ly 4096 --steps 20000 --sample 100]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Dict, Iterable, Sequence

import numpy as np

from . import Channel, Lattice

SWEEP = {"f64": "panel6", "f32": "panel4", "f16": "panel4"}
CHECKPOINTS = (1000, 2000, 5000, 10000)


def velocity_error(ref, got) -> Dict[str, float]:
    du = np.concatenate([(got[1] - ref[1]).ravel(), (got[2] - ref[2]).ravel()])
    u = np.concatenate([ref[1].ravel(), ref[2].ravel()])
    return {"u_l2_rel": float(np.linalg.norm(du) / np.linalg.norm(u)),
            "u_max_rel": float(np.abs(du).max() / np.abs(u).max())}


def series_error(ref: np.ndarray, got: np.ndarray) -> Dict[str, float]:
    out = {}
    for k, name in ((0, "drag"), (1, "lift")):
        scale = max(float(np.abs(ref[:, k]).max()), 1e-300)
        out[f"{name}_max_rel"] = float(np.abs(got[:, k] - ref[:, k]).max() / scale)
        out[f"{name}_scale"] = scale
    # the open channel's mass changes in FP64 too (inlet and outlet): the
    # error is the difference from the FP64 series; the drift is the run's own
    m, m64 = got[:, 3], ref[:, 3]
    out["mass_max_rel"] = float(np.abs(m - m64).max() / np.abs(m64).max())
    out["mass_drift_rel"] = float(np.abs(m - m[0]).max() / m[0])
    return out


def lockstep(ch: Channel, steps: int, sample: int, storages: Sequence[str] = ("f64", "f32", "f16"),
             checkpoints: Iterable[int] = CHECKPOINTS, sweep: Dict[str, str] = None,
             log=lambda m: None) -> dict:
    """Advance one channel in every storage form together; returns per form
    the drag/lift/mass series, the final fields and the velocity errors
    against FP64 at each checkpoint and at the end."""
    sweep = dict(SWEEP, **(sweep or {}))
    lats = {s: Lattice(ch, storage=s) for s in storages}
    try:
        for lat in lats.values():
            lat.init_equilibrium()
        series = {s: [lat.diagnostics()] for s, lat in lats.items()}
        seconds = {s: 0.0 for s in storages}
        errs = {s: [] for s in storages if s != "f64"}
        marks = sorted(set(c for c in checkpoints if 0 < c < steps))
        done = 0
        while done < steps:
            n = min(sample, steps - done)
            for s, lat in lats.items():
                t0 = time.perf_counter()
                lat.step(n, sweep[s])
                lat.sync()
                seconds[s] += time.perf_counter() - t0
                series[s].append(lat.diagnostics())
            done += n
            if done in marks:
                ref = lats["f64"].macros()
                for s in errs:
                    errs[s].append(dict(velocity_error(ref, lats[s].macros()), steps=done))
                log(f"step {done}: " + ", ".join(f"{s} {errs[s][-1]['u_l2_rel']:.3g}"
                                                 for s in errs))
        fields = {s: lat.macros() for s, lat in lats.items()}
    finally:
        for lat in lats.values():
            lat.close()
    for s in errs:
        errs[s].append(dict(velocity_error(fields["f64"], fields[s]), steps=steps))
    return {"series": {s: np.array(v) for s, v in series.items()}, "fields": fields,
            "errors": errs, "seconds": seconds, "sweep": {s: sweep[s] for s in storages}}


def study(lx: int = 16384, ly: int = 4096, reynolds: Iterable[float] = (100.0, 1000.0),
          steps: int = 20000, sample: int = 100, storages=("f32", "f16"), log=print) -> dict:
    result = {"lattice": [lx, ly], "disc": None, "steps": steps, "sample_every": sample,
              "sweep": {s: SWEEP[s] for s in ("f64",) + tuple(storages)}, "runs": []}
    for re in reynolds:
        ch = Channel.geometry(lx, ly, "B", reynolds=float(re))
        result["disc"] = [ch.obst_x, ch.obst_y, ch.obst_r]
        r = lockstep(ch, steps, sample, ("f64",) + tuple(storages), log=log)
        ref = r["series"]["f64"]
        m = ref[:, 3]
        entry = {"reynolds": float(re), "omega": ch.omega(),
                 "f64": {"seconds": round(r["seconds"]["f64"], 2),
                         "mass_drift_rel": float(np.abs(m - m[0]).max() / m[0]),
                         "drag_final": float(ref[-1, 0]), "lift_final": float(ref[-1, 1]),
                         "drag_scale": float(np.abs(ref[:, 0]).max()),
                         "lift_scale": float(np.abs(ref[:, 1]).max())}}
        for s in storages:
            final = r["errors"][s][-1]
            entry[s] = dict(u_l2_rel=final["u_l2_rel"], u_max_rel=final["u_max_rel"],
                            **series_error(ref, r["series"][s]),
                            seconds=round(r["seconds"][s], 2),
                            u_l2_rel_by_steps={str(e["steps"]): e["u_l2_rel"]
                                               for e in r["errors"][s]})
            log(f"Re {re}: {s} {entry[s]}")
        result["runs"].append(entry)
        del r
    return result


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--lx", type=int, default=16384)
    ap.add_argument("--ly", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--sample", type=int, default=100)
    ap.add_argument("--reynolds", default="100,1000")
    ap.add_argument("--out", default=None, help="write the JSON result here as well")
    a = ap.parse_args(argv)
    res = study(a.lx, a.ly, [float(r) for r in a.reynolds.split(",")], a.steps, a.sample,
                log=lambda m: print(m, flush=True))
    res["command"] = "cd memory-aware-lbm_amd && python3 -m lbm_amd.precision " + " ".join(
        sys.argv[1:] if argv is None else argv)
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
