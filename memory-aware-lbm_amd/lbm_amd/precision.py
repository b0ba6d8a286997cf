"""Storage-precision study: FP32 and FP16 deviation storage against the
bit-exact FP64 path on the same channel.

Workload: the 16384 x 4096 cylinder channel, disc (lx/5, ly/2, ly/10), at
Re 100 and Re 1000, 20000 steps each.  Every storage form runs the panel
sweep (two steps per launch) from the same initial state.  Drag and lift
(momentum exchange on the disc, FP64 reduction) are sampled every `sample`
steps; at the end the velocity fields are compared.

Reported per (Re, storage):
  * velocity: L2 relative error ||u - u64|| / ||u64|| over both components,
    and max relative error max|u - u64| / max|u64|;
  * drag / lift series: max |d - d64| / max |d64| (same for lift);
  * mass drift of each run relative to its own start.

The FP64 run is the reference of the comparison; it equals the reference C
code bit for bit at 256 x 128 (tests/test_gpu_parity.py), and a CPU run of
this size and length (1.3e12 lattice updates) is out of reach.

    python -m lbm_amd.precision [--lx 16384 --ly 4096 --steps 20000 --sample 100]
"""
from __future__ import annotations

import argparse
import json
import time
from typing import Dict, Iterable

import numpy as np

from . import Channel, Lattice


def run(ch: Channel, storage: str, steps: int, sample: int, variant: str = "panel") -> dict:
    """Advance one channel `steps` steps; return diagnostics samples and the
    final velocity field."""
    series = []
    t0 = time.perf_counter()
    with Lattice(ch, storage=storage) as lat:
        lat.init_equilibrium()
        series.append(lat.diagnostics())
        done = 0
        while done < steps:
            n = min(sample, steps - done)
            lat.step(n, variant)
            done += n
            series.append(lat.diagnostics())
        lat.sync()
        _, ux, uy = lat.macros()
    return {"diag": np.array(series), "ux": ux, "uy": uy,
            "seconds": time.perf_counter() - t0}


def compare(ref: dict, got: dict) -> Dict[str, float]:
    du = np.concatenate([(got["ux"] - ref["ux"]).ravel(), (got["uy"] - ref["uy"]).ravel()])
    u = np.concatenate([ref["ux"].ravel(), ref["uy"].ravel()])
    out = {"u_l2_rel": float(np.linalg.norm(du) / np.linalg.norm(u)),
           "u_max_rel": float(np.abs(du).max() / np.abs(u).max())}
    for k, name in ((0, "drag"), (1, "lift")):
        d = got["diag"][:, k] - ref["diag"][:, k]
        out[f"{name}_max_rel"] = float(np.abs(d).max() / max(np.abs(ref["diag"][:, k]).max(),
                                                               1e-300))
    m = got["diag"][:, 3]
    out["mass_drift_rel"] = float(np.abs(m - m[0]).max() / m[0])
    return out


def study(lx: int = 16384, ly: int = 4096, reynolds: Iterable[float] = (100.0, 1000.0),
          steps: int = 20000, sample: int = 100, storages=("f32", "f16"),
          variant: str = "panel", log=print) -> dict:
    result = {"lattice": [lx, ly], "steps": steps, "sample_every": sample,
              "variant": variant, "runs": []}
    for re in reynolds:
        ch = Channel.geometry(lx, ly, "B", reynolds=float(re))
        ref = run(ch, "f64", steps, sample, variant)
        m = ref["diag"][:, 3]
        entry = {"reynolds": float(re), "omega": ch.omega(),
                 "f64": {"seconds": round(ref["seconds"], 2),
                         "mass_drift_rel": float(np.abs(m - m[0]).max() / m[0]),
                         "drag_final": float(ref["diag"][-1, 0]),
                         "lift_final": float(ref["diag"][-1, 1])}}
        log(f"Re {re}: f64 done in {ref['seconds']:.1f} s")
        for s in storages:
            got = run(ch, s, steps, sample, variant)
            entry[s] = dict(compare(ref, got), seconds=round(got["seconds"], 2))
            log(f"Re {re}: {s} {entry[s]}")
            del got
        del ref
        result["runs"].append(entry)
    return result


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--lx", type=int, default=16384)
    ap.add_argument("--ly", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--sample", type=int, default=100)
    ap.add_argument("--reynolds", default="100,1000")
    ap.add_argument("--variant", default="panel")
    ap.add_argument("--out", default=None, help="write the JSON result here as well")
    a = ap.parse_args(argv)
    res = study(a.lx, a.ly, [float(r) for r in a.reynolds.split(",")], a.steps, a.sample,
                variant=a.variant, log=lambda m: print(m, flush=True))
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
