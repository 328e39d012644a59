"""GPU tier: batched scenarios in one concatenated population, graph replay
of the learned form, and the autotuner.

Scenarios are independent populations; concatenating them with per-scenario
offsets must give, per scenario, exactly what a run of that scenario alone
gives (states bit for bit, the same events with indices shifted by the
scenario offset). Graph replay and autotuned launch shapes must not change
any result.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPOCH = 2461253.9166666665
SPECS = [(4000, 11, 2), (9000, 12, 0), (6000, 13, 3)]  # (catalogue size, seed, breakups)


@pytest.fixture(scope="module")
def scenarios(kamd):
    pops = [kamd.cascade(n, seed, ev, 0.1) for n, seed, ev in SPECS]
    off = np.concatenate([[0], np.cumsum([len(p) for p in pops])])
    return pops, off


@pytest.fixture(scope="module")
def net(kamd):
    eng = kamd.Engine(0)
    eng.set_time(EPOCH, 0.0, 60.0)
    eng.net_train(kamd.generate(200000, seed=101)["state"], samples_per_orbit=15, spacing_steps=12,
                  iterations=3000, batch=16384, seed=1)
    p, norm = eng.net_get()
    eng.close()
    return p, norm


def _run(eng, pop, steps, dt, precision, offsets=None, graph=False, thr=5.0):
    eng.set_precision(precision)
    eng.set_graph(graph)
    eng.set_objects(pop["state"], pop["alive"].astype(np.uint8), pop["size_m"])
    if offsets is not None:
        eng.set_scenarios(offsets)
    eng.set_config(thr, 1.0, 0.01)
    eng.set_time(EPOCH, 0.0, dt)
    eng.run(steps, screen=True)
    s, a = eng.get_objects()
    return s, a, eng.take_events(), eng.take_stats()


def _key(ev, shift=0):
    return sorted((int(e["i"]) - shift, int(e["j"]) - shift, int(e["step"]), float(e["t"]), float(e["miss_km"]),
                   float(e["v_rel_km_s"]), int(e["collision"])) for e in ev)


STAT_FIELDS = ("candidate_pairs", "refined_pairs", "co_orbiting", "cell_km", "max_cell_objects", "cell_doublings",
               "propagated")  # dense_candidates depends on the tiling by design


def _same_stats(a, b):
    return all(np.array_equal(a[f], b[f]) for f in STAT_FIELDS)


def _compare_split(pops, off, s, a, ev, singles):
    sid = np.searchsorted(off, ev["i"], side="right") - 1
    sjd = np.searchsorted(off, ev["j"], side="right") - 1
    assert np.array_equal(sid, sjd), "a pair crosses scenarios"
    for k, (s1, a1, e1, _) in enumerate(singles):
        lo, hi = off[k], off[k + 1]
        assert np.array_equal(s[lo:hi], s1), f"scenario {k} states"
        assert np.array_equal(a[lo:hi], a1), f"scenario {k} alive"
        assert _key(ev[sid == k], lo) == _key(e1), f"scenario {k} events"


def test_concatenated_plain_matches_separate_runs(engine, kamd, scenarios):
    pops, off = scenarios
    allpop = np.concatenate(pops)
    s, a, ev, _ = _run(engine, allpop, 60, 10.0, kamd.PRECISION_FP64, offsets=off)
    singles = [_run(engine, p, 60, 10.0, kamd.PRECISION_FP64) for p in pops]
    print(f"plain: {len(allpop)} objects in {len(pops)} scenarios, {len(ev)} events "
          f"({[len(x[2]) for x in singles]} per scenario alone)")
    assert len(ev) > 100
    _compare_split(pops, off, s, a, ev, singles)


def test_concatenated_learned_matches_separate_runs_and_graph(engine, kamd, scenarios, net):
    pops, off = scenarios
    engine.net_set(*net)
    allpop = np.concatenate(pops)
    s, a, ev, _ = _run(engine, allpop, 30, 60.0, kamd.PRECISION_LEARNED, offsets=off)
    singles = [_run(engine, p, 30, 60.0, kamd.PRECISION_LEARNED) for p in pops]
    assert len(ev) > 100
    _compare_split(pops, off, s, a, ev, singles)
    # The same run replayed from captured chunk graphs. With chunks of 5
    # steps the chunks after the first few have one shape (length, check
    # pattern), so one captured graph serves all of them.
    engine.set_chunk(5)
    try:
        c0, l0 = engine.graph_counts()
        sg, ag, evg, _ = _run(engine, allpop, 30, 60.0, kamd.PRECISION_LEARNED, offsets=off, graph=True)
        c1, l1 = engine.graph_counts()
    finally:
        engine.set_chunk(256)
        engine.set_graph(False)
    print(f"learned: {len(ev)} events; graph captures {c1 - c0}, launches {l1 - l0}")
    assert l1 - l0 >= 8 and 1 <= c1 - c0 < l1 - l0
    assert np.array_equal(sg, s) and np.array_equal(ag, a)
    assert _key(evg) == _key(ev)


def test_autotune_keeps_results(engine, kamd, scenarios):
    pops, off = scenarios
    allpop = np.concatenate(pops)
    ref = _run(engine, allpop, 20, 10.0, kamd.PRECISION_FP64, offsets=off)
    engine.set_precision(kamd.PRECISION_FP64)
    engine.set_objects(allpop["state"], allpop["alive"].astype(np.uint8), allpop["size_m"])
    engine.set_scenarios(off)
    engine.set_config(5.0, 1.0, 0.01)
    engine.set_time(EPOCH, 0.0, 10.0)
    best, base_ms, best_ms = engine.autotune(steps=4)
    print(f"autotune: {best}, {base_ms:.2f} ms -> {best_ms:.2f} ms for 4 steps")
    assert best_ms <= base_ms * 1.05
    assert engine.time() == (0.0, 0)
    engine.run(20, screen=True)
    s, a = engine.get_objects()
    ev, st = engine.take_events(), engine.take_stats()
    assert np.array_equal(s, ref[0]) and np.array_equal(a, ref[1])
    assert _key(ev) == _key(ref[2])
    assert _same_stats(st, ref[3])
    # Deliberately different tilings give the same results too: tiny dense
    # cells and short ranges; and large sparse cells (more objects than a
    # wave has lanes) whose long ranges go to work items chunk by chunk.
    # Both sparse-cell kernels.
    for tiling in (dict(dense_min=8, long_range=16, seg=64, dense_blocks=300, exact_blocks=100, sparse_kernel=1),
                   dict(dense_min=160, long_range=96, seg=128, dense_blocks=512, exact_blocks=256, sparse_kernel=0),
                   dict(dense_min=160, long_range=96, seg=128, dense_blocks=512, exact_blocks=256, sparse_kernel=1)):
        engine.set_tuning(**tiling)
        s2, a2, ev2, st2 = _run(engine, allpop, 20, 10.0, kamd.PRECISION_FP64, offsets=off)
        assert np.array_equal(s2, ref[0]) and _key(ev2) == _key(ref[2]) and _same_stats(st2, ref[3]), tiling
    engine.set_tuning(dense_min=128, long_range=256, seg=256, dense_blocks=1024, exact_blocks=512, cell_mult=1.0,
                      sparse_kernel=0)
