"""GPU tier: khip_screen_static, the allocation-free screening entry point,
called through the C ABI and compared with khip_screen on the same inputs:
events (every field) and statistics bit for bit, on two calls in a row, the
second with a smaller population (the fixed device array keeps the first
call's contents, so everything the second call reads must be cleared or
rewritten by it). Also: the context's own population is left untouched, and
the limits are reported as errors, not faults.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPOCH = 2461253.9166666665


def _cloud(n, seed):
    """n objects near one 550 km orbit (phase, plane and speed spread): dense
    enough for thousands of conjunctions per 10 s step."""
    rng = np.random.default_rng(seed)
    r0 = 6378.137 + 550.0
    v0 = (398600.4418 / r0) ** 0.5
    u = rng.uniform(0, 0.05, n)
    inc = np.radians(53.0 + rng.normal(0, 0.3, n))
    raan = rng.uniform(0, 0.05, n)
    cu, su, ci, si, co, so = np.cos(u), np.sin(u), np.cos(inc), np.sin(inc), np.cos(raan), np.sin(raan)
    s = np.zeros((n, 6))
    s[:, 0] = (co * cu - so * su * ci) * r0
    s[:, 1] = (so * cu + co * su * ci) * r0
    s[:, 2] = su * si * r0
    s[:, 3] = (-co * su - so * cu * ci) * v0
    s[:, 4] = (-so * su + co * cu * ci) * v0
    s[:, 5] = cu * si * v0
    s[:, 3:] *= rng.normal(1.0, 1e-3, (n, 1))
    return s


def _population(kamd, n_cat, n_cloud, seed):
    cat = kamd.generate(n_cat, seed=seed)
    before = np.concatenate([cat["state"], _cloud(n_cloud, seed)])
    size = np.concatenate([cat["size_m"], np.full(n_cloud, 0.8)])
    alive = np.ones(len(before), dtype=np.uint8)
    alive[::97] = 0
    return before, alive, size


def _after(kamd, before, size):
    eng = kamd.Engine(0)
    try:
        eng.set_objects(before, np.ones(len(before), np.uint8), size)
        eng.set_time(EPOCH, 0.0, 10.0)
        eng.run(1, screen=False)
        after, _ = eng.get_objects()
    finally:
        eng.close()
    return after


def _static(kamd, eng, before, after, alive, size, t0, dt, cap):
    lib = kamd.library()
    before = np.ascontiguousarray(before, dtype=np.float64)
    after = np.ascontiguousarray(after, dtype=np.float64)
    alive = np.ascontiguousarray(alive, dtype=np.uint8)
    size = np.ascontiguousarray(size, dtype=np.float64)
    out = np.zeros(max(cap, 1), dtype=kamd.EVENT_DTYPE)
    st = np.zeros(1, dtype=kamd.STATS_DTYPE)
    ne = C.c_int64(-1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib.khip_screen_static(eng.h, len(before), p(before), p(after), p(alive), p(size), float(t0), float(dt),
                                p(out), cap, C.byref(ne), p(st))
    return rc, out[: max(ne.value, 0)].copy() if rc == 0 else None, st[0], ne.value


def _same(ev_a, st_a, ev_b, st_b):
    assert len(ev_a) == len(ev_b)
    for f in ("i", "j", "step", "t", "miss_km", "v_rel_km_s", "collision"):
        assert np.array_equal(ev_a[f], ev_b[f]), f
    for f in ("candidate_pairs", "refined_pairs", "co_orbiting", "cell_km", "max_cell_objects", "cell_doublings",
              "survivors"):
        assert st_a[f] == st_b[f], f


def test_static_screening_matches_khip_screen_twice(kamd):
    big = _population(kamd, 26000, 6000, 21)
    small = _population(kamd, 3000, 1500, 22)
    eng = kamd.Engine(0)
    try:
        eng.set_config(5.0, 1.0, 0.01)
        mine = kamd.generate(1000, seed=5)
        eng.set_objects(mine["state"], mine["alive"].astype(np.uint8), mine["size_m"])
        eng.set_time(EPOCH, 0.0, 10.0)
        held = eng.get_objects()
        for k, (before, alive, size) in enumerate((big, small)):
            after = _after(kamd, before, size)
            t0 = 3600.0 * (k + 1)
            rc, ev_s, st_s, ne = _static(kamd, eng, before, after, alive, size, t0, 10.0, 1 << 17)
            assert rc == 0, kamd.library().khip_last_error(eng.h)
            ev_r, st_r = eng.screen(before, after, alive, size, t0, 10.0)
            print(f"call {k + 1}: {len(before)} objects, {ne} events, {int(st_s['candidate_pairs'])} candidates")
            assert ne == len(ev_s) and len(ev_s) > (300 if k == 0 else 20)
            _same(ev_s, st_s, ev_r, st_r)
            # khip_screen overwrote the context's population; set it again
            # and check that the static call leaves it alone.
            eng.set_objects(mine["state"], mine["alive"].astype(np.uint8), mine["size_m"])
            _static(kamd, eng, before, after, alive, size, t0, 10.0, 1 << 17)
            s, a = eng.get_objects()
            assert np.array_equal(s, held[0]) and np.array_equal(a, held[1])
    finally:
        eng.close()


def test_static_screening_limits(kamd):
    eng = kamd.Engine(0)
    try:
        eng.set_config(5.0, 1.0, 0.01)
        before, alive, size = _population(kamd, 1000, 3000, 23)
        after = _after(kamd, before, size)
        rc, _, _, ne = _static(kamd, eng, before, after, alive, size, 0.0, 10.0, 1 << 17)
        assert rc == 0 and ne > 10
        rc, _, _, ne2 = _static(kamd, eng, before, after, alive, size, 0.0, 10.0, 5)
        assert rc == -3 and ne2 == ne  # buffer too small: count reported, nothing copied
        many = np.zeros((40000, 6))
        many[:, 0] = 7000.0
        many[:, 4] = 7.5
        rc, _, _, _ = _static(kamd, eng, many, many, np.ones(40000, np.uint8), np.full(40000, 0.8), 0.0, 10.0, 16)
        assert rc == -3
        assert b"at most" in kamd.library().khip_last_error(eng.h)
    finally:
        eng.close()
