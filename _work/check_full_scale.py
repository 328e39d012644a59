"""Full-scale parity check of the plain FP64 path (not collected by pytest:
three minutes of CPU on 16 threads for the cascade, about fourteen for the
1M-object catalogue).

Populations: `cascade`, the 28,340-object synthetic catalogue of seed 0
plus 40 seeded catastrophic breakups at t = 0 (255,220 objects), screened
at 5 km; `million`, 1,000,000 synthetic objects (seed 0, LEO shells x1.5),
screened at 2 km. Either is stepped for 24 h at dt = 10 s (radius scale 1,
v_min 10 m/s) twice: by the engine's plain FP64 path on the GPU, and by the
reference's own compiled CPU path (oracle/_ref, the main.cpp step loop:
Simulation::step + Screener). Both runs go in ten-minute chunks, the reference
chained through its returned states with t_start = chunk start (its CPU
screener keeps no state between steps). Checked: the event lists, field by
field and bit for bit (pair, step, TCA, miss distance, relative speed,
collision flag, in (step, t, i, j) order), the final states bit for bit and
the alive flags. Prints one JSON line.

    python tests/check_full_scale.py [--population cascade|million] [--hours H]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402
from oracle.bind import reference  # noqa: E402

EPOCH = 2461253.916666667
DT = 10.0
CHUNK = 60  # steps per chunk (ten minutes)
FIELDS = ("i", "j", "step", "t", "miss_km", "v_rel_km_s", "collision")


def ordered(ev):
    return ev[np.lexsort((ev["j"], ev["i"], ev["t"], ev["step"]))]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--hours", type=int, default=24)
    ap.add_argument("--population", choices=("cascade", "million"), default="cascade")
    args = ap.parse_args()
    nchunks = args.hours * 3600 // int(CHUNK * DT)
    nsteps = nchunks * CHUNK
    kamd = load_package()
    if args.population == "cascade":
        pop, thr = kamd.cascade(28340, 0, 40, 0.045), 5.0
    else:
        pop, thr = kamd.generate(1_000_000, seed=0, mix=(0.7, 0.1, 0.1, 0.1, 1.5)), 2.0
    alive0 = pop["alive"].astype(np.uint8)
    print(f"{len(pop)} objects, {nsteps} steps", flush=True)

    eng = kamd.Engine(0)
    eng.set_precision(0)
    eng.set_objects(pop["state"], alive0, pop["size_m"])
    eng.set_config(thr, 1.0, 0.01)
    eng.set_time(EPOCH, 0.0, DT)
    t1 = time.perf_counter()
    gev = []
    for c in range(nchunks):
        eng.run(CHUNK, screen=True)
        gev.append(eng.take_events())
        eng.take_stats()
    gpu_s = time.perf_counter() - t1
    gs, ga = eng.get_objects()
    eng.close()
    gev = ordered(np.concatenate(gev))
    print(f"engine: {len(gev)} events in {gpu_s:.1f} s", flush=True)

    ref = reference()
    if ref is None:
        raise SystemExit("oracle/_ref is not built")
    cores = ref.threads(0)
    states, alive = pop["state"].copy(), alive0.copy()
    rev, cpu_s = [], 0.0
    for c in range(nchunks):
        states, alive, ev, _, loop_s = ref.run(states, alive, pop["size_m"], EPOCH, DT, CHUNK, screen=True, thr=thr,
                                               t_start=c * CHUNK * DT, cap=1 << 23, per_step=False)
        ev = ev.copy()
        ev["step"] += c * CHUNK
        rev.append(ev)
        cpu_s += loop_s
        print(f"reference: chunk {c + 1}/{nchunks}, {len(ev)} events, {loop_s:.1f} s", flush=True)
    rev = ordered(np.concatenate(rev))

    same_len = len(gev) == len(rev)
    fields = {f: bool(same_len and np.array_equal(gev[f], rev[f])) for f in FIELDS}
    out = {
        "population": args.population, "objects": int(len(pop)), "threshold_km": thr, "steps": nsteps,
        "events_engine": int(len(gev)), "events_reference": int(len(rev)),
        "collisions": int(gev["collision"].sum()), "fields_equal": fields,
        "states_equal": bool(np.array_equal(gs, states, equal_nan=True)),
        "alive_equal": bool(np.array_equal(ga, alive)),
        "engine_s": gpu_s, "reference_loop_s": cpu_s, "reference_threads": cores,
    }
    # Rows that differ (NaN in the same place counts as equal), by alive flag.
    diff = ~((gs == states) | (np.isnan(gs) & np.isnan(states))).all(axis=1)
    out["rows_differing"] = {"alive": int((diff & (alive == 1)).sum()), "dead": int((diff & (alive == 0)).sum())}
    out["rows_with_nan"] = {"engine": int(np.isnan(gs).any(axis=1).sum()), "reference": int(np.isnan(states).any(axis=1).sum())}
    print(json.dumps(out), flush=True)
    ok = same_len and all(fields.values()) and out["states_equal"] and out["alive_equal"]
    raise SystemExit(0 if ok else 1)


if __name__ == "__main__":
    main()
