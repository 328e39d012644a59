"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in a container that has the reference checkout, after
``make -C oracle ref`` built oracle/_ref/libkessler_ref.so from the
reference's own sources. Every expected output below comes from the
reference's code (through oracle/ref_harness.cpp), never from this
repository's oracle or engine.

Fixtures (all data):
  catalog_20260801_1000Z.psv.gz  the reference's real catalogue
      (data/satellites_20260801_1000Z.txt), reduced to the fields the
      loader reads (orbital.cpp:150-164): id, type, RCS class and the six
      state strings, verbatim; header kept for the epoch line.
  ref_vectors.npz   unit vectors: moon_position, acceleration, rk4,
      filter_pair, prefilter_keep, interpolate_state, screen_step scenes,
      PortableRng draws.
  ref_catalog_1h.npz  1 h (360 steps, dt 10 s, 5 km) main-loop run on the
      real catalogue: events, per-step stats, final states.
  ref_breakup.npz   Fragmentation::impact in sequence on small synthetic
      catalogues (inputs from the package's generator, stored with the
      outputs): every object after the impacts and the per-event
      fragment / escaped counts (the cases of tests/test_breakup_host.py).

``python tests/golden/make_golden.py --breakup`` writes ref_breakup.npz only.
"""
from __future__ import annotations

import gzip
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.bind import reference  # noqa: E402

REF_DATA = "/root/reference/data/satellites_20260801_1000Z.txt"
EPOCH = 2461253.9166666665


def reduce_catalogue(src: str, dst: str) -> None:
    out = []
    with open(src, encoding="utf-8", errors="replace") as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line.startswith("#"):
                out.append(line)
                continue
            fld = [x.strip() for x in line.split("|")]
            if len(fld) < 17:
                continue
            keep = [fld[0], "", fld[2], "", "", fld[5], "", "", "", "", ""] + fld[11:17]
            out.append("|".join(keep))
    with gzip.open(dst, "wt", encoding="utf-8", compresslevel=9) as g:
        g.write("\n".join(out) + "\n")


def random_states(rng, n):
    """Orbit-like states from LEO to beyond GEO."""
    r = rng.uniform(6600.0, 45000.0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = rng.normal(size=(n, 3))
    w -= (w * u).sum(1, keepdims=True) * u
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    vc = np.sqrt(398600.4418 / r) * rng.uniform(0.8, 1.2, n)
    s = np.zeros((n, 6))
    s[:, :3] = u * r[:, None]
    s[:, 3:] = w * vc[:, None] + rng.normal(scale=0.05, size=(n, 3))
    return s


def relative_pairs(rng, n, dt):
    """Relative states around the filter's decision boundaries."""
    r = rng.normal(size=(n, 3))
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    r *= np.where(rng.random(n) < 0.5, rng.uniform(0, 6, n), rng.uniform(0, 200, n))[:, None]
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    speed = np.where(rng.random(n) < 0.2, rng.uniform(0, 0.02, n), rng.uniform(0, 15, n))
    v *= speed[:, None]
    return r, v


def screen_scene(rng, n, spread_km, speed_km_s):
    """A tight cloud: start-of-step and end-of-step states of n objects."""
    base = random_states(rng, 1)[0]
    before = np.repeat(base[None, :], n, axis=0)
    before[:, :3] += rng.normal(scale=spread_km, size=(n, 3))
    before[:, 3:] += rng.normal(scale=speed_km_s, size=(n, 3))
    return before


def main() -> None:
    ref = reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libkessler_ref.so not built: run make -C oracle ref")
    ref.threads(8)
    rng = np.random.default_rng(20260801)

    reduce_catalogue(REF_DATA, os.path.join(HERE, "catalog_20260801_1000Z.psv.gz"))

    g = {}
    # moon_position at 128 Julian days over two centuries.
    jd = np.concatenate([EPOCH + rng.uniform(-36525, 36525, 120),
                         EPOCH + np.arange(8) * (10.0 / 86400.0)])
    g["moon_jd"] = jd
    g["moon_pos"] = np.array([ref.moon_position(x) for x in jd])

    # acceleration and rk4 on 1024 states, Moon from the table above.
    st = random_states(rng, 1024)
    mi = rng.integers(0, len(jd), size=(1024, 3))
    g["acc_r"] = st[:, :3]
    g["acc_rm"] = g["moon_pos"][mi[:, 0]]
    g["acc_out"] = np.array([ref.acceleration(a, b) for a, b in zip(g["acc_r"], g["acc_rm"])])
    g["rk4_s"] = st
    g["rk4_dt"] = rng.choice([1.0, 10.0, 30.0, 60.0, 120.0], 1024)
    g["rk4_m"] = g["moon_pos"][mi]
    g["rk4_out"] = np.array([ref.rk4(s, d, m[0], m[1], m[2])
                             for s, d, m in zip(st, g["rk4_dt"], g["rk4_m"])])

    # filter_pair on 20k relative states; prefilter_keep on their FP32 rounding.
    r, v = relative_pairs(rng, 20000, 10.0)
    dts = rng.choice([10.0, 30.0, 60.0], 20000)
    thr = rng.choice([2.0, 5.0], 20000)
    g["fp_r"], g["fp_v"], g["fp_dt"], g["fp_thr"] = r, v, dts, thr
    g["fp_out"] = np.array([ref.filter_pair(a, b, d, t, 1e-4) for a, b, d, t in zip(r, v, dts, thr)],
                           dtype=np.int32)
    g["pk_out"] = np.array([ref.prefilter_keep(a.astype(np.float32), b.astype(np.float32), d, t, 1e-4)
                            for a, b, d, t in zip(r, v, dts, thr)], dtype=np.int32)

    # interpolate_state (Hermite over a step) on 256 (a, b, dt, s).
    a = random_states(rng, 256)
    b = a.copy()
    b[:, :3] += a[:, 3:] * 10.0
    s = rng.uniform(0, 1, 256)
    g["hi_a"], g["hi_b"], g["hi_s"] = a, b, s
    g["hi_out"] = np.array([ref.interpolate_state(x, y, 10.0, z) for x, y, z in zip(a, b, s)])

    # screen_step on small clouds: before = cloud, after = one exact RK4 step.
    scenes = []
    for k, (n, spread, speed) in enumerate([(60, 3.0, 0.5), (200, 8.0, 2.0), (400, 30.0, 7.0),
                                            (150, 0.5, 0.005), (300, 15.0, 0.2)]):
        before = screen_scene(rng, n, spread, speed)
        m = g["moon_pos"][120:123]
        after = np.array([ref.rk4(x, 10.0, m[0], m[1], m[2]) for x in before])
        alive = (rng.random(n) > 0.05).astype(np.uint8)
        size_m = rng.choice([0.2, 0.8, 4.0, 200.0], n)
        ev, stats = ref.screen(before, after, alive, size_m, 3600.0, 10.0, 5.0, 1.0, 0.01)
        scenes.append((before, after, alive, size_m, ev, stats))
        g[f"sc{k}_before"], g[f"sc{k}_after"], g[f"sc{k}_alive"] = before, after, alive
        g[f"sc{k}_size"] = size_m
        g[f"sc{k}_events"] = ev
        g[f"sc{k}_stats"] = np.array([stats])
    g["n_scenes"] = np.array(len(scenes))

    # PortableRng (breakup.cpp:121-160) and the raw mt19937_64 stream.
    g["rng_uniform"] = ref.rng(1, 0, 1000)
    g["rng_normal"] = ref.rng(1, 1, 1000)
    g["rng_poisson3"] = ref.rng(1, 2, 1000, mean=3.0)
    g["rng_poisson1500"] = ref.rng(1, 2, 1000, mean=1500.0)
    g["rng_raw_default"] = ref.rng(5489, 4, 10000)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **g)

    # 1 h of the main loop on the real catalogue.
    cat = ref.load_catalog(REF_DATA)
    n = cat["states"].shape[0]
    alive = np.ones(n, dtype=np.uint8)
    fs, fa, ev, stats, _ = ref.run(cat["states"], alive, cat["size_m"], cat["epoch_jd"], 10.0, 360)
    idx = np.sort(rng.choice(n, 1024, replace=False))
    np.savez_compressed(os.path.join(HERE, "ref_catalog_1h.npz"),
                        events=ev, stats=stats, sample_idx=idx, sample_states=fs[idx],
                        final_sha256=np.array(hashlib.sha256(fs.tobytes()).hexdigest()),
                        alive=fa, epoch_jd=np.array(cat["epoch_jd"]))
    print("golden fixtures written:", sorted(os.listdir(HERE)))


def breakup_fixture() -> None:
    ref = reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libkessler_ref.so not built: run make -C oracle ref")
    from tests.test_breakup_host import CASES, _case
    from __graft_entry__ import load_package

    kamd = load_package()
    g = {}
    for k in range(len(CASES)):
        pop, targets, pm, vr, seed, lc = _case(kamd, k)
        objs, ev = ref.impacts(pop["state"], pop["size_m"], pop["mass_kg"], EPOCH, 0.0, targets, pm, vr, seed, lc)
        g[f"in_state_{k}"], g[f"in_size_{k}"], g[f"in_mass_{k}"] = pop["state"], pop["size_m"], pop["mass_kg"]
        g[f"targets_{k}"], g[f"pm_{k}"], g[f"vr_{k}"] = targets, pm, vr
        g[f"seed_{k}"], g[f"lc_{k}"] = np.array(seed), np.array(lc)
        n0 = len(pop)
        # Parents keep their states; only their alive flags and masses change.
        g[f"out_alive_{k}"], g[f"out_mass_kg_{k}"] = objs["alive"], objs["mass_kg"]
        for f in ("event", "state", "size_m", "area_to_mass"):
            g[f"out_{f}_{k}"] = objs[f][n0:]
        g[f"out_events_{k}"] = np.stack([ev["fragments"], ev["escaped"]], 1)
    g["n_cases"] = np.array(len(CASES))
    np.savez_compressed(os.path.join(HERE, "ref_breakup.npz"), **g)
    print("ref_breakup.npz written")


if __name__ == "__main__":
    if "--breakup" in sys.argv:
        breakup_fixture()
    else:
        main()
        breakup_fixture()
