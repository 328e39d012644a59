"""GPU tier: the HIP engine, called through the C ABI, against the oracle and
the compiled reference's golden fixtures.

Parity bar: the same event set (pair, step), TCA within 1 ms,
miss within 1 m, final states within 1e-9 relative. The engine's FP64 path
keeps the reference's operation order without contraction, so every check
here asks for more: bit-identical states, events and per-step statistics.
"""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPOCH = 2461253.9166666665
EVENT_FIELDS = ("i", "j", "step", "t", "miss_km", "v_rel_km_s", "collision")
STAT_FIELDS = ("candidate_pairs", "refined_pairs", "co_orbiting", "cell_km", "max_cell_objects")


def _same_events(got, exp, fields=EVENT_FIELDS):
    assert len(got) == len(exp), (len(got), len(exp))
    for f in fields:
        assert np.array_equal(got[f], exp[f]), f


def _same_stats(got, exp):
    assert len(got) == len(exp)
    for f in STAT_FIELDS:
        assert np.array_equal(got[f], exp[f]), f


def _run_engine(engine, states, alive, size_m, dt, steps, thr=5.0, chunk=None, epoch=EPOCH):
    engine.set_objects(states, alive, size_m)
    engine.set_config(thr, 1.0, 0.01)
    engine.set_time(epoch, 0.0, dt)
    if chunk:
        engine.set_chunk(chunk)
    engine.run(steps, screen=True)
    s, a = engine.get_objects()
    return s, a, engine.take_events(), engine.take_stats()


def test_fp64_division_sqrt_floor_are_ieee(engine):
    """The bit-identical claim rests on correctly rounded FP64 div and sqrt."""
    rng = np.random.default_rng(11)
    n = 1 << 20
    a = np.concatenate([rng.uniform(0, 1e6, n // 4), np.exp(rng.uniform(-40, 40, n // 4)),
                        rng.uniform(-5e4, 5e4, n // 4), rng.normal(size=n // 4) ** 2])
    b = np.concatenate([rng.uniform(1, 700, n // 4), np.exp(rng.uniform(-40, 40, n // 4)),
                        rng.uniform(150, 260, n // 4), rng.uniform(0.5, 2, n // 4)])
    out = engine.fp64_selfcheck(a, b)
    with np.errstate(invalid="ignore"):
        assert np.array_equal(out[:, 0], np.sqrt(a), equal_nan=True)
    assert np.array_equal(out[:, 1], a / b)
    assert np.array_equal(out[:, 2], np.floor(a / b))


@pytest.mark.parametrize("k", range(5))
def test_screen_scenes_match_reference(engine, golden, k):
    ev, st = engine.screen(golden[f"sc{k}_before"], golden[f"sc{k}_after"], golden[f"sc{k}_alive"],
                           golden[f"sc{k}_size"], 3600.0, 10.0)
    _same_events(ev, golden[f"sc{k}_events"], EVENT_FIELDS[:2] + EVENT_FIELDS[3:])
    est = golden[f"sc{k}_stats"][0]
    for f in STAT_FIELDS:
        assert st[f] == est[f], f


def test_catalog_one_hour_matches_reference(engine, catalog, golden_1h):
    n = catalog["states"].shape[0]
    s, a, ev, st = _run_engine(engine, catalog["states"], np.ones(n, np.uint8), catalog["size_m"],
                               10.0, 360)
    assert hashlib.sha256(s.tobytes()).hexdigest() == str(golden_1h["final_sha256"])
    assert np.array_equal(a, golden_1h["alive"])
    _same_events(ev, golden_1h["events"])
    _same_stats(st, golden_1h["stats"])


def test_graph_replay_matches_reference_hour(engine, catalog, golden_1h):
    """The plain path's chunks captured into HIP graphs and replayed (chunks
    of 16 steps: one capture serves every full chunk) give the reference's
    hour bit for bit, as direct launches do."""
    n = catalog["states"].shape[0]
    engine.set_graph(True)
    try:
        c0, l0 = engine.graph_counts()
        s, a, ev, st = _run_engine(engine, catalog["states"], np.ones(n, np.uint8), catalog["size_m"], 10.0, 360,
                                   chunk=16)
        c1, l1 = engine.graph_counts()
    finally:
        engine.set_graph(False)
        engine.set_chunk(256)
    print(f"graph captures {c1 - c0}, launches {l1 - l0}")
    assert l1 - l0 >= 20 and 1 <= c1 - c0 <= 10
    assert hashlib.sha256(s.tobytes()).hexdigest() == str(golden_1h["final_sha256"])
    assert np.array_equal(a, golden_1h["alive"])
    _same_events(ev, golden_1h["events"])
    _same_stats(st, golden_1h["stats"])


def test_chunking_does_not_change_results(engine, catalog, orc):
    n = 4000
    states = catalog["states"][:n]
    alive = np.ones(n, np.uint8)
    size_m = catalog["size_m"][:n]
    s1, a1, e1, t1 = _run_engine(engine, states, alive, size_m, 10.0, 50, chunk=7)
    s2, a2, e2, st2, _ = orc.run(states, alive, size_m, EPOCH, 10.0, 50)
    assert np.array_equal(s1, s2) and np.array_equal(a1, a2)
    _same_events(e1, e2)
    _same_stats(t1, st2)


def _cloud(rng, n, spread_km, speed_km_s, r0=6928.137):
    v0 = np.sqrt(398600.4418 / r0)
    s = np.zeros((n, 6))
    s[:, 0] = r0
    s[:, 4] = v0
    s[:, :3] += rng.normal(scale=spread_km, size=(n, 3))
    s[:, 3:] += rng.normal(scale=speed_km_s, size=(n, 3))
    return s


def test_dense_cloud_overflows_event_buffer(engine, orc):
    """More than 65,536 events in one step: the engine's re-run path."""
    rng = np.random.default_rng(12)
    n = 1500
    before = _cloud(rng, n, 1.5, 0.3)
    # Converging on the cloud centre at 0.1 /s: most pairs cross mid-step.
    before[:, 3:] -= (before[:, :3] - before[:, :3].mean(axis=0)) * 0.1
    after = np.array([orc.rk4(x, 10.0, [1e15, 0, 0], [1e15, 0, 0], [1e15, 0, 0]) for x in before])
    alive = np.ones(n, np.uint8)
    size_m = np.full(n, 0.8)
    exp, est = orc.screen(before, after, alive, size_m, 0.0, 10.0, thr=5.0)
    assert len(exp) > 65536, len(exp)
    got, st = engine.screen(before, after, alive, size_m, 0.0, 10.0, cap=len(exp) + 10)
    _same_events(got, exp, EVENT_FIELDS[:2] + EVENT_FIELDS[3:])
    for f in STAT_FIELDS:
        assert st[f] == est[f], f


def test_reentries_dead_objects_and_fresh_fragments(engine, orc):
    """Decaying orbits (alive flags flip mid-run), objects dead from the
    start, and a co-located fragment cloud (one cell with 800 objects)."""
    rng = np.random.default_rng(13)
    parts = [_cloud(rng, 1000, 300.0, 0.5)]
    low = _cloud(rng, 200, 20.0, 0.1, r0=6378.137 + 80.0)
    low[:, 3:] *= 0.8  # sub-orbital: re-enter within minutes
    parts.append(low)
    frag = np.repeat(_cloud(rng, 1, 0.0, 0.0), 800, axis=0)
    frag[:, 3:] += rng.normal(scale=0.1, size=(800, 3))
    parts.append(frag)
    states = np.concatenate(parts)
    n = states.shape[0]
    alive = np.ones(n, np.uint8)
    alive[rng.choice(n, 50, replace=False)] = 0
    size_m = rng.choice([0.2, 0.8, 4.0], n)
    s1, a1, e1, t1 = _run_engine(engine, states, alive, size_m, 10.0, 60)
    s2, a2, e2, t2, _ = orc.run(states, alive, size_m, EPOCH, 10.0, 60)
    assert (a2 == 0).sum() > 100, "the scene must contain re-entries"
    assert np.array_equal(s1, s2) and np.array_equal(a1, a2)
    _same_events(e1, e2)
    _same_stats(t1, t2)
    assert t1["max_cell_objects"].max() >= 750  # the cloud, less the objects marked dead


def test_empty_and_single_object(engine, orc):
    s, a, e, t = _run_engine(engine, np.zeros((0, 6)), np.zeros(0, np.uint8), np.zeros(0), 10.0, 3)
    assert len(e) == 0 and len(t) == 3 and np.all(t["candidate_pairs"] == 0)
    one = np.array([[7000.0, 0, 0, 0, 7.546, 0]])
    s, a, e, t = _run_engine(engine, one, np.ones(1, np.uint8), np.ones(1), 10.0, 5)
    s2, a2, e2, t2, _ = orc.run(one, np.ones(1, np.uint8), np.ones(1), EPOCH, 10.0, 5)
    assert np.array_equal(s, s2) and len(e) == 0
    _same_stats(t, t2)


def test_catalog_24h_known_answer(engine, catalog):
    """README.md:177-187: the real catalogue over 24 h at dt = 10 s, 5 km:
    38,367 conjunctions, 1,493 under 1 km, 14 under 100 m, no collision,
    closest 7.2 m between NORAD 49157 and 64849 (7.17728 m at 1.6221 h)."""
    n = catalog["states"].shape[0]
    s, a, ev, st = _run_engine(engine, catalog["states"], np.ones(n, np.uint8), catalog["size_m"],
                               10.0, 8640)
    assert len(ev) == 38367
    assert (ev["miss_km"] < 1.0).sum() == 1493
    assert (ev["miss_km"] < 0.1).sum() == 14
    assert (ev["miss_km"] < 2.0).sum() == 5451
    assert ev["collision"].sum() == 0
    k = np.argmin(ev["miss_km"])
    ids = {int(catalog["ids"][ev["i"][k]]), int(catalog["ids"][ev["j"][k]])}
    assert ids == {49157, 64849}
    assert abs(ev["miss_km"][k] * 1000 - 7.17728) < 5e-5
    assert abs(ev["t"][k] / 3600 - 1.6221) < 5e-4
    assert len(st) == 8640
