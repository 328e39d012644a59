"""GPU tier: in-run breakups on the device (khip_set_breakup), in the plain
form and in the learned form (60 s steps; collisions confirmed by the plain
re-verification; fragments start on the plain method).

Scene: pairs of LEO objects set on head-on-ish courses (relative speed
10 km/s) that meet half-way through the first 10 s step, far from each
other; radius scale 1000 so every meeting is a collision, and every
collision is catastrophic (500 m_p v^2 / m_t = 25,000 J/g >> 40). For the
distributions: 40 pairs of 10 t / 5 t parents 0.5 m across, so the
fragments' mass stays far below the budget and the drawn size law is what
is kept. For determinism and chunking: 200 pairs of 1 t / 0.5 t parents.

The device draws come from a counter-based generator, so the fragments are
checked against the model's closed forms rather than against the
reference's exact draws:
  * fragment count per event ~ Poisson(0.1 M^0.75 L_c^-1.71), M = 1,500 kg;
  * sizes follow the power law truncated to [L_c, parent size]
    (Kolmogorov-Smirnov against the exact CDF);
  * log10(delta-v [m/s]) - (0.9 log10(A/m) + 2.9) ~ N(0, 0.4);
  * fragment mass never exceeds the parents' mass; both parents die;
  * fragments are screened from the next step on; results do not depend on
    chunking or on how often the reserved object room had to grow.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPOCH = 2461253.9166666665
MU = 398600.4418
LC = 0.1
SLOPE = 1.71
PARENT_L = 0.5


def _scene(npairs=200, seed=3, masses_kg=(1000.0, 500.0)):
    rng = np.random.default_rng(seed)
    states, sizes, masses = [], [], []
    for _ in range(npairs):
        rhat = rng.normal(size=3)
        rhat /= np.linalg.norm(rhat)
        r0 = rhat * (6378.137 + 600.0)
        vc = math.sqrt(MU / np.linalg.norm(r0))
        a = np.cross(rhat, rng.normal(size=3))
        a /= np.linalg.norm(a)
        vt = a * vc
        theta = 2 * math.asin(10.0 / (2 * vc))  # |vp - vt| = 10 km/s
        b = np.cross(rhat, a)
        vp = (a * math.cos(theta) + b * math.sin(theta)) * vc
        for v, size, m in ((vt, PARENT_L, masses_kg[0]), (vp, PARENT_L, masses_kg[1])):
            states.append(np.concatenate([r0 - v * 5.0, v]))  # meet at t = 5 s
            sizes.append(size)
            masses.append(m)
    return np.array(states), np.array(sizes), np.array(masses)


def _run(engine, steps=3, chunk=256, reserve=65536, seed=9, npairs=200, masses_kg=(1000.0, 500.0), precision=0,
         dt=10.0, graph=False):
    s, size, mass = _scene(npairs, masses_kg=masses_kg)
    engine.set_precision(precision)
    engine.set_chunk(chunk)
    engine.set_graph(graph)
    engine.set_objects(s, np.ones(len(s), np.uint8), size)
    engine.set_masses(mass)
    engine.set_config(5.0, 1000.0, 0.01)
    engine.set_time(EPOCH, 0.0, dt)
    engine.set_breakup(True, reserve=reserve, lc_m=LC, seed=seed)
    try:
        engine.run(steps, screen=True)
    finally:
        engine.set_breakup(False)
        engine.set_chunk(256)
        engine.set_graph(False)
    states, alive = engine.get_objects()
    return (states, alive, engine.take_events(), engine.take_breakups(), engine.sizes(), engine.masses(),
            engine.take_stats())


def _ks(sample, cdf):
    x = np.sort(sample)
    f = cdf(x)
    n = len(x)
    return max(np.max(np.arange(1, n + 1) / n - f), np.max(f - np.arange(n) / n))


def test_breakups_follow_the_model(engine):
    npairs = 40
    n0 = 2 * npairs
    states, alive, ev, bk, size, mass, _ = _run(engine, steps=1, npairs=npairs, masses_kg=(10000.0, 5000.0))
    coll = ev[ev["collision"] == 1]
    print(f"{len(bk)} breakups, {int(bk['fragments'].sum())} fragments, {int(bk['escaped'].sum())} escaped; "
          f"collisions in the step: {len(coll)}")
    assert len(bk) == npairs and len(coll) >= npairs
    assert (bk["catastrophic"] == 1).all()
    assert np.allclose(bk["energy_j_per_g"], 500 * 5000 * 100 / 10000, rtol=1e-3)
    # Both parents of every event die.
    assert not alive[:n0].any()
    # Counts: Poisson(expected).
    expected = 0.1 * 15000 ** 0.75 * LC ** -SLOPE
    assert np.allclose(bk["expected_fragments"], expected, rtol=1e-12)
    z = (bk["drawn"] - expected) / math.sqrt(expected)
    print(f"drawn per event: mean z {z.mean():.3f}, var {z.var():.3f}")
    assert abs(z.mean()) < 0.6 and 0.4 < z.var() < 1.8
    # Mass budget and bookkeeping.
    assert (bk["fragments_mass_kg"] <= bk["sbm_mass_kg"] + 1e-9).all()
    nfrag = int(bk["fragments"].sum())
    assert len(states) == n0 + nfrag
    order = np.argsort(bk["first_index"])
    starts = bk["first_index"][order]
    assert starts[0] == n0 and np.array_equal(np.diff(starts), bk["fragments"][order][:-1])
    frag = slice(n0, n0 + nfrag)
    assert alive[frag].mean() > 0.99
    # Sizes: the power law truncated to [L_c, parent size].
    L = size[frag]
    assert (L >= LC).all() and (L <= PARENT_L).all()
    assert (bk["fragments_mass_kg"] < 0.5 * bk["sbm_mass_kg"]).all()  # the budget never binds here
    cdf = lambda x: (1 - (x / LC) ** -SLOPE) / (1 - (PARENT_L / LC) ** -SLOPE)
    d = _ks(L, cdf)
    print(f"sizes: KS distance {d:.4f} over {len(L)} fragments")
    assert d < 1.63 / math.sqrt(len(L))  # 1% level
    # Delta-v law: fragment velocity minus its parent's at the end of the step
    # (the parents are frozen there; same place, same gravity). The parent is
    # the one of the event whose velocity is closer.
    ev_of = order[np.searchsorted(starts, np.arange(n0, n0 + nfrag), side="right") - 1]
    vt = states[bk["target"][ev_of], 3:]
    vp = states[bk["projectile"][ev_of], 3:]
    vf = states[frag, 3:]
    near_t = np.linalg.norm(vf - vt, axis=1) < np.linalg.norm(vf - vp, axis=1)
    dv = np.linalg.norm(vf - np.where(near_t[:, None], vt, vp), axis=1) * 1e3
    chi = np.log10(0.540424 * L * L / mass[frag])
    res = np.log10(dv) - (0.9 * chi + 2.9)
    print(f"delta-v residual: mean {res.mean():.4f}, std {res.std():.4f}")
    assert abs(res.mean()) < 0.02 and abs(res.std() - 0.4) < 0.02


def test_breakups_are_deterministic_and_chunk_independent(engine):
    a = _run(engine, steps=4, chunk=256, reserve=65536)
    b = _run(engine, steps=4, chunk=1, reserve=1000)  # the reserved room grows several times
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert len(a[2]) == len(b[2]) and np.array_equal(np.sort(a[2], order=["step", "t", "i", "j"]),
                                                     np.sort(b[2], order=["step", "t", "i", "j"]))
    assert np.array_equal(a[3], b[3])
    assert np.array_equal(a[6]["candidate_pairs"], b[6]["candidate_pairs"])
    # Fragments take part in screening from the step after their birth: the
    # siblings of each cloud are candidate pairs (and recede, so the exact
    # filter rejects them).
    st = a[6]
    print(f"candidate pairs per step: {st['candidate_pairs'].tolist()}, refined {st['refined_pairs'].tolist()}")
    assert st["candidate_pairs"][0] < 10000 and st["candidate_pairs"][1] > 1_000_000


def test_learned_form_breakups_match_plain(engine, trained):
    """The same 40 collisions in the learned form (one 60 s step, which is a
    check step: every object takes the plain method) and in the plain form
    (six 10 s steps): the same collisions in the same order give the same
    breakup records, the same fragment draws (sizes, masses), and fragment
    states within the learned form's 10 m bound (parents at the collision
    from the 60 s step's cubic, whose velocity is off by about a cm/s a few
    seconds into the step; fragments stepped to the end of the 60 s step in
    one RK4 step)."""
    _, _, params, norm = trained
    engine.net_set(params, norm)
    kw = dict(npairs=40, masses_kg=(10000.0, 5000.0))
    sp, ap, evp, bkp, szp, msp, _ = _run(engine, steps=6, **kw)
    sl, al, evl, bkl, szl, msl, stl = _run(engine, steps=2, precision=2, dt=60.0, **kw)
    print(f"plain: {len(bkp)} breakups, {int(bkp['fragments'].sum())} fragments; learned: {len(bkl)} breakups, "
          f"{int(bkl['fragments'].sum())} fragments")
    assert len(bkp) == len(bkl) == 40
    for f in ("id", "target", "projectile", "catastrophic", "drawn", "fragments", "escaped", "first_index"):
        assert np.array_equal(bkp[f], bkl[f]), f
    assert np.allclose(bkp["expected_fragments"], bkl["expected_fragments"], rtol=1e-12)
    assert np.allclose(bkp["t"], bkl["t"], atol=1e-6)
    assert len(sp) == len(sl)
    n0 = 80
    assert np.array_equal(szp[n0:], szl[n0:]) and np.array_equal(msp[n0:], msl[n0:])
    # The states at one instant: 120 s, two learned steps and twelve plain ones.
    sp2, ap2, _, _, _, _, _ = _run(engine, steps=12, **kw)
    print(f"alive after 120 s: plain {int(ap2.sum())}, learned {int(al.sum())}, differing {int((ap2 != al).sum())}")
    assert np.array_equal(ap2, al)
    both = (ap2 == 1) & (al == 1)
    d = np.linalg.norm(sp2[both, :3] - sl[both, :3], axis=1) * 1e3
    print(f"fragment positions after 120 s: |dr| max {d.max():.4f} m, median {np.median(d):.5f} m")
    assert d.max() < 10.0
    # The fragments are screened from the second step on (cloud siblings).
    assert stl["candidate_pairs"][6:].sum() > 100 * max(1, stl["candidate_pairs"][:6].sum())


def test_learned_form_breakups_deterministic_and_chunk_independent(engine, trained):
    _, _, params, norm = trained
    engine.net_set(params, norm)
    a = _run(engine, steps=4, chunk=256, reserve=65536, precision=2, dt=60.0)
    b = _run(engine, steps=4, chunk=1, reserve=1000, precision=2, dt=60.0)  # the room grows several times
    print(f"learned: {len(a[3])} breakups, {int(a[3]['fragments'].sum())} fragments, {len(a[2])} events")
    assert len(a[3]) >= 200  # the 200 meetings, then collisions inside the clouds (radius scale 1000)
    rows = np.flatnonzero(~(a[0] == b[0]).all(axis=1))
    if len(rows):
        first = np.searchsorted(np.sort(a[3]["first_index"]), rows, side="right")
        print(f"rows differing: {len(rows)} (alive a/b: {int(a[1][rows].sum())}/{int(b[1][rows].sum())}; "
              f"first {rows[:8].tolist()}; breakup ordinal {first[:8].tolist()}; max |dr| "
              f"{np.abs(a[0][rows, :3] - b[0][rows, :3]).max():.3e} km)")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert len(a[2]) == len(b[2]) and np.array_equal(np.sort(a[2], order=["step", "t", "i", "j"]),
                                                     np.sort(b[2], order=["step", "t", "i", "j"]))
    assert np.array_equal(a[3], b[3])
    assert np.array_equal(a[6]["candidate_pairs"], b[6]["candidate_pairs"])


@pytest.mark.parametrize("precision,dt", [(0, 10.0), (2, 60.0)])
def test_breakups_in_captured_graphs_equal_eager(engine, trained, precision, dt):
    """Chunks with in-run breakups replayed from captured graphs (the step's
    time from the step table) give the eager run's states, events and
    breakup records bit for bit, in the plain and the learned form."""
    _, _, params, norm = trained
    engine.net_set(params, norm)
    a = _run(engine, steps=8, chunk=2, reserve=65536, precision=precision, dt=dt)
    c0, l0 = engine.graph_counts()
    b = _run(engine, steps=8, chunk=2, reserve=65536, precision=precision, dt=dt, graph=True)
    c1, l1 = engine.graph_counts()
    print(f"precision {precision}: {len(a[3])} breakups, {len(a[2])} events; graph captures {c1 - c0}, "
          f"launches {l1 - l0}")
    assert l1 - l0 >= 4 and len(a[3]) > 0
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2].tobytes() == b[2].tobytes() and np.array_equal(a[3], b[3])
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])
