"""Calibration of the CPU baseline: the oracle's restatement of the step loop
and the reference's compiled loop (oracle/_ref) on the 28,340-object
synthetic catalogue (seed 0, the catalogue orbit mix), 24 h at dt = 10 s,
screening at 5 km (radius scale 1, v_min 10 m/s), on all host threads.
Prints one JSON line: bitwise agreement (events, final states, alive flags)
and both loop times.

    python oracle/calibrate.py [--hours H]
"""
import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402
from oracle.bind import oracle, reference  # noqa: E402

EPOCH = 2461253.916666667


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--hours", type=float, default=24.0)
    args = ap.parse_args()
    steps = int(round(args.hours * 360))
    pop = load_package().generate(28340, seed=0)
    alive = pop["alive"].astype(np.uint8)
    orc, ref = oracle(), reference()
    if ref is None:
        raise SystemExit("oracle/_ref is not built")
    threads = orc.threads(0)
    ref.threads(0)
    out = {}
    runs = {}
    for name, lib in (("oracle", orc), ("reference", ref)):
        runs[name] = lib.run(pop["state"], alive, pop["size_m"], EPOCH, 10.0, steps, screen=True, thr=5.0,
                             cap=1 << 24, per_step=False)
        out[f"{name}_loop_s"] = runs[name][4]
    (so, ao, eo, _, _), (sr, ar, er, _, _) = runs["oracle"], runs["reference"]
    out.update({"objects": int(len(pop)), "steps": steps, "threads": threads, "events": int(len(er)),
                "events_equal": bool(len(eo) == len(er) and all(np.array_equal(eo[f], er[f]) for f in er.dtype.names)),
                "states_equal": bool(np.array_equal(so, sr)), "alive_equal": bool(np.array_equal(ao, ar)),
                "oracle_over_reference_time": out["oracle_loop_s"] / out["reference_loop_s"]})
    print(json.dumps(out))


if __name__ == "__main__":
    main()
