"""Benchmark of the MI355X propagate-and-screen engine (one GPU).

Workload: the cascade-scale plain-FP64 configuration, a 28,340-object
synthetic catalogue (orbit mix of the reference's 2026-08-01 catalogue,
seed 0) plus 40 seeded catastrophic NASA-breakup collisions at t = 0
(target and projectile 100-2000 kg, 8-12 km/s, L_c = 0.045 m), 255,220
objects in all; FP64 RK4 (Earth + J2 + Moon) at dt = 10 s with screening
at 5 km (radius scale 1, minimum encounter speed 10 m/s). A step is one
simulation step of the whole population: propagation plus screening.

value = object-steps/s = (sum over the timed steps of the objects propagated)
/ wall time of the timed steps, with the population already resident in HBM.

The JSON line also carries the roofline of the dominant kernel (HIP-event
timing on the engine's stream during the timed steps) and the CPU baseline:
the reference's own OpenMP CPU path (oracle/_ref, compiled from the
reference sources) timed on this host over a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

METRIC = "object-steps/s incl. screening, learned/mixed vs plain FP64 at <=10 m/24h"
EPOCH = 2461253.916666667
DT = 10.0
THRESHOLD_KM = 5.0
# Algorithmic work per unit: FP32 flops per candidate pair of
# the pre-filter (prefilter_keep + building dr, dv), HBM bytes per propagated
# object of the RK4 step (FP64 state in and out, alive flag).
FLOP_PER_CANDIDATE = 60.0
BYTES_PER_OBJECT_STEP = 97.0
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3


def cascade_population(kamd):
    return kamd.cascade(28340, 0, 40, 0.045)


def run_gpu(kamd, pop, steps, warmup, timing=True):
    eng = kamd.Engine(0)
    alive = pop["alive"].astype(np.uint8)

    def reset():
        eng.set_objects(pop["state"], alive, pop["size_m"])
        eng.set_config(THRESHOLD_KM, 1.0, 0.01)
        eng.set_time(EPOCH, 0.0, DT)

    if warmup > 0:
        reset()
        eng.run(warmup, screen=True)
        eng.take_events()
        eng.take_stats()
    reset()
    eng.set_timing(timing)
    eng.reset_timing()
    t0 = time.perf_counter()
    eng.run(steps, screen=True)
    wall = time.perf_counter() - t0
    tm = eng.timing()
    ev = eng.take_events()
    st = eng.take_stats()
    return eng, wall, tm, ev, st


def cpu_baseline(kamd, eng, pop, windows_h=(0.0, 6.0, 12.0, 18.0), budget_s=24.0):
    """The compiled reference CPU path over a bounded sample: a few steps from
    each of several start times of the same run. The start states come from
    the engine, which is bit-identical to the CPU path."""
    from oracle.bind import oracle, reference

    ref = reference()
    kind = "reference"
    if ref is None:
        ref, kind = oracle(), "port"
    cores = ref.threads(0)
    alive0 = pop["alive"].astype(np.uint8)
    size_m = pop["size_m"]
    per_window = budget_s / len(windows_h)
    obj_steps, cpu_s, steps_done, sample = 0, 0.0, 0, []
    for h in windows_h:
        nstart = int(round(h * 3600.0 / DT))
        eng.set_objects(pop["state"], alive0, size_m)
        eng.set_config(THRESHOLD_KM, 1.0, 0.01)
        eng.set_time(EPOCH, 0.0, DT)
        if nstart:
            eng.run(nstart, screen=False)
        states, alive = eng.get_objects()
        t_start = nstart * DT
        spent, k = 0.0, 0
        while spent < per_window and k < 50:
            n_alive = int(alive.sum())
            states, alive, _, _, loop_s = ref.run(states, alive, size_m, EPOCH, DT, 1, screen=True,
                                                  thr=THRESHOLD_KM, t_start=t_start + k * DT,
                                                  per_step=False)
            spent += loop_s
            obj_steps += n_alive
            k += 1
        cpu_s += spent
        steps_done += k
        sample.append(f"{k} steps from t={h:g} h")
    return {
        "value": obj_steps / cpu_s,
        "unit": "object-steps/s",
        "cores": cores,
        "kind": kind,
        "sample": (f"{steps_done} steps of the same workload ({'; '.join(sample)}), start states "
                   f"from the engine (bit-identical to the CPU path), {cpu_s:.1f} s of CPU loop time"),
    }


def match_events(e_ref, e_new):
    """Recall of e_ref's (pair, step) set in e_new, and the TCA / miss spread of the matches."""
    ref = {(int(a), int(b), int(s)): (t, m) for a, b, s, t, m in
           zip(e_ref["i"], e_ref["j"], e_ref["step"], e_ref["t"], e_ref["miss_km"])}
    new = {(int(a), int(b), int(s)): (t, m) for a, b, s, t, m in
           zip(e_new["i"], e_new["j"], e_new["step"], e_new["t"], e_new["miss_km"])}
    common = set(ref) & set(new)
    dt = max((abs(ref[k][0] - new[k][0]) for k in common), default=0.0)
    dm = max((abs(ref[k][1] - new[k][1]) for k in common), default=0.0)
    return {"reference_events": len(ref), "events": len(new), "matched": len(common),
            "recall": len(common) / max(1, len(ref)), "max_dtca_ms": dt * 1e3, "max_dmiss_m": dm * 1e3}


def precision_workload(kamd, args):
    """1,000,000 objects (LEO shells densified), dt 10 s, 24 h,
    screening at 2 km; the precision form (FP32 J2 + Moon, FP64 state and
    central term) against the plain FP64 path on the same catalogue."""
    thr = 2.0
    pop = kamd.generate(1_000_000, seed=0, mix=(0.7, 0.1, 0.1, 0.1, 1.5))
    eng = kamd.Engine(0)
    res = {}
    for name, mode in (("fp64", kamd.PRECISION_FP64), ("mixed", kamd.PRECISION_MIXED)):
        eng.set_precision(mode)
        eng.set_objects(pop["state"], pop["alive"].astype(np.uint8), pop["size_m"])
        eng.set_config(thr, 1.0, 0.01)
        eng.set_time(EPOCH, 0.0, DT)
        if args.warmup:
            eng.run(args.warmup, screen=True)
            eng.take_events()
            eng.take_stats()
            eng.set_objects(pop["state"], pop["alive"].astype(np.uint8), pop["size_m"])
            eng.set_time(EPOCH, 0.0, DT)
        eng.set_timing(True)
        eng.reset_timing()
        t0 = time.perf_counter()
        eng.run(args.steps, screen=True)
        wall = time.perf_counter() - t0
        st = eng.take_stats()
        res[name] = dict(wall=wall, events=eng.take_events(), stats=st, timing=eng.timing(),
                         states=eng.get_objects())
    s64, a64 = res["fp64"]["states"]
    s32, a32 = res["mixed"]["states"]
    both = (a64 == 1) & (a32 == 1)
    dr = np.linalg.norm(s32[both, :3] - s64[both, :3], axis=1) * 1e3
    dv = np.linalg.norm(s32[both, 3:] - s64[both, 3:], axis=1) * 1e6
    obj_steps = int(res["mixed"]["stats"]["propagated"].sum())
    v_mixed = obj_steps / res["mixed"]["wall"]
    v_fp64 = int(res["fp64"]["stats"]["propagated"].sum()) / res["fp64"]["wall"]
    tm = res["mixed"]["timing"]
    out = {
        "metric": METRIC, "value": v_mixed, "unit": "object-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * res["mixed"]["wall"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 state + f32 perturbations",
        "data": "synthetic",
        "config": {"workload": "precision form: 1,000,000 synthetic objects (seed 0, LEO shells x1.5), "
                               f"dt 10 s, {args.steps * DT / 3600:g} h, screening 2 km",
                   "objects": int(len(pop)), "dt_s": DT, "threshold_km": thr},
        "fp64_value": v_fp64, "mixed_over_fp64": v_mixed / v_fp64,
        "position_error_m": {"max": float(dr.max()), "p99": float(np.percentile(dr, 99)),
                             "median": float(np.median(dr))},
        "velocity_error_mm_s": {"max": float(dv.max()), "p99": float(np.percentile(dv, 99))},
        "alive_flags_equal": bool(np.array_equal(a64, a32)),
        "events_vs_fp64": match_events(res["fp64"]["events"], res["mixed"]["events"]),
        "candidate_pairs_per_step": float(res["mixed"]["stats"]["candidate_pairs"].mean()),
        "kernel_ms_per_step": {k: float(tm[k]) / max(1, int(tm["steps"])) for k in
                               ("propagate_ms", "grid_ms", "pairs_sparse_ms", "pairs_dense_ms",
                                "exact_ms")},
        "device": eng.device_name,
    }
    eng.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8640)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="cascade", choices=("cascade", "precision"),
                    help="cascade: configs[1] (default, the headline line); precision: configs[2]")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: --gpus must be 1 "
                                   "(multi-GPU sharding is a design in DESIGN.md, not built)"}))
        sys.exit(2)

    kamd = load_package()
    if args.workload == "precision":
        precision_workload(kamd, args)
        return
    pop = cascade_population(kamd)
    eng, wall, tm, ev, st = run_gpu(kamd, pop, args.steps, args.warmup)
    obj_steps = int(st["propagated"].sum())
    cand = int(st["candidate_pairs"].sum())
    value = obj_steps / wall

    dense_cand = int(st["dense_candidates"].sum())
    kernels = {"k_pairs": float(tm["pairs_sparse_ms"]), "k_pairs_dense": float(tm["pairs_dense_ms"]),
               "k_exact": float(tm["exact_ms"]), "grid": float(tm["grid_ms"]),
               "k_propagate": float(tm["propagate_ms"])}
    single = {k: v for k, v in kernels.items() if k != "grid"}
    dominant = max(single, key=single.get)
    launches = max(1, int(tm["steps"]))

    def valu_roof(name, ms, cand_count):
        achieved = FLOP_PER_CANDIDATE * cand_count / (ms * 1e-3) / 1e12
        return {"kernel": name, "bound": "valu-fp32", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                "avg_launch_us": 1e3 * ms / launches, "units_per_launch": cand_count / launches,
                "unit_of_work": "candidate pair (60 FP32 flops)"}

    prop_gbs = BYTES_PER_OBJECT_STEP * obj_steps / (tm["propagate_ms"] * 1e-3) / 1e9
    roof_prop = {"kernel": "k_propagate", "bound": "hbm", "achieved": prop_gbs, "peak": PEAK_HBM_GBS,
                 "unit": "GB/s", "frac": prop_gbs / PEAK_HBM_GBS, "traffic": None,
                 "avg_launch_us": 1e3 * tm["propagate_ms"] / launches,
                 "units_per_launch": obj_steps / launches,
                 "unit_of_work": "propagated object (97 B: FP64 state in + out, alive flag)"}
    if dominant == "k_pairs_dense":
        roof = valu_roof("k_pairs_dense", tm["pairs_dense_ms"], dense_cand)
    elif dominant == "k_pairs":
        roof = valu_roof("k_pairs", tm["pairs_sparse_ms"], cand - dense_cand)
    elif dominant == "k_propagate":
        roof = roof_prop
    else:
        roof = {"kernel": "k_exact", "bound": "valu-fp64", "achieved": None, "peak": 78.6,
                "unit": "TFLOP/s", "frac": None, "traffic": None,
                "avg_launch_us": 1e3 * tm["exact_ms"] / launches}
    roof["kernel_ms_per_step"] = {k: v / launches for k, v in kernels.items()}
    roof["reruns"] = int(tm["reruns"])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "object-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"cascade: 28,340-object synthetic catalogue (seed 0) + 40 catastrophic "
                         f"breakups at t=0 (L_c 0.045 m) = {len(pop):,} objects; plain FP64 RK4 "
                         f"Earth+J2+Moon, dt 10 s, screening 5 km; steps 0..{args.steps}"),
            "objects": int(len(pop)),
            "dt_s": DT,
            "threshold_km": THRESHOLD_KM,
            "simulated_h": args.steps * DT / 3600.0,
        },
        "object_seconds_per_s": value * DT,
        "events": int(len(ev)),
        "collisions": int(ev["collision"].sum()),
        "candidate_pairs": cand,
        "roofline": roof,
        "roofline_propagate": roof_prop,
        "device": eng.device_name,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(kamd, eng, pop)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
