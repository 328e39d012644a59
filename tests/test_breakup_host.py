"""CPU tier: the host NASA breakup model (host/breakup.cpp, reached through
khost_impacts / khost_cascade) against the compiled reference's
Fragmentation::impact (breakup.cpp:121-360, through oracle/_ref).

A seed must give the reference's fragments bit for bit: same count, sizes,
masses, area-to-mass ratios, birth states and alive flags, and the same
per-event summaries. Cases: catastrophic impacts (Poisson means above 10:
the PTRS branch), a non-catastrophic one (small mean: the product branch and
the surviving target's mass), a parent smaller than L_c, and a sequence of
impacts sharing one generator. The committed fixture
(tests/golden/ref_breakup.npz, written from the reference by make_golden.py)
pins the same outputs where oracle/_ref is absent.
"""
import os

import numpy as np
import pytest

EPOCH = 2461253.9166666665
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_breakup.npz")

# (catalogue seed, targets chosen from LARGE LEO objects, projectile kg,
#  impact speed km/s, generator seed, L_c m)
CASES = [
    (4, 3, (500.0, 80.0, 1500.0), (10.0, 11.5, 8.0), 7, 0.05),
    (5, 1, (0.05,), (3.0,), 11, 0.05),       # non-catastrophic, Poisson mean 9.2
    (6, 2, (300.0, 2.0), (9.0, 12.0), 3, 0.2),
]


def _case(kamd, k):
    cat_seed, ntarget, pm, vr, seed, lc = CASES[k]
    pop = kamd.generate(2000, seed=cat_seed)
    r = np.linalg.norm(pop["state"][:, :3], axis=1)
    pool = np.flatnonzero((pop["size_m"] >= 4.0) & (r < 6378.137 + 2100.0))
    targets = pool[:ntarget].astype(np.int32)
    if k == 2:  # one parent below L_c: a 0.2 m object hit by a 2 kg projectile
        targets[1] = np.flatnonzero((pop["size_m"] == 0.2) & (r < 6378.137 + 2100.0))[0]
    return pop, targets, np.array(pm), np.array(vr), seed, lc


@pytest.mark.parametrize("k", range(len(CASES)))
def test_host_impacts_match_reference(kamd, ref, k):
    if ref is None:
        pytest.skip("compiled reference (oracle/_ref) not built here")
    pop, targets, pm, vr, seed, lc = _case(kamd, k)
    mine, ev = kamd.impacts(pop, EPOCH, 0.0, targets, pm, vr, seed, lc)
    theirs, rev = ref.impacts(pop["state"], pop["size_m"], pop["mass_kg"], EPOCH, 0.0, targets, pm, vr, seed, lc)
    print(f"case {k}: {len(mine) - len(pop)} fragments, events {ev.tolist()}")
    assert len(mine) == len(theirs)
    for f in ("event", "alive", "state", "size_m", "mass_kg", "area_to_mass"):
        assert np.array_equal(mine[f], theirs[f]), f
    assert (mine["type"][len(pop):] == 1).all() and np.array_equal(theirs["type"][len(pop):], mine["type"][len(pop):])
    assert np.array_equal(ev[:, 0], rev["fragments"]) and np.array_equal(ev[:, 1], rev["escaped"])
    if k == 1:
        assert not rev["catastrophic"].any()
        # The surviving target lost the ejecta mass, as in the reference.
        assert mine["alive"][targets[0]] == 1 and mine["mass_kg"][targets[0]] < pop["mass_kg"][targets[0]]


def test_host_impacts_match_golden(kamd):
    g = np.load(GOLDEN)
    for k in range(int(g["n_cases"])):
        n = len(g[f"in_state_{k}"])
        pop = np.zeros(n, dtype=kamd.OBJECT_DTYPE)
        pop["id"] = np.arange(1, n + 1)
        pop["event"] = -1
        pop["alive"] = 1
        pop["state"] = g[f"in_state_{k}"]
        pop["size_m"] = g[f"in_size_{k}"]
        pop["mass_kg"] = g[f"in_mass_{k}"]
        mine, ev = kamd.impacts(pop, EPOCH, 0.0, g[f"targets_{k}"], g[f"pm_{k}"], g[f"vr_{k}"],
                                int(g[f"seed_{k}"]), float(g[f"lc_{k}"]))
        assert len(mine) == len(g[f"out_alive_{k}"]), k
        for f in ("alive", "mass_kg"):
            assert np.array_equal(mine[f], g[f"out_{f}_{k}"]), (k, f)
        for f in ("event", "state", "size_m", "area_to_mass"):
            assert np.array_equal(mine[f][n:], g[f"out_{f}_{k}"]), (k, f)
        assert np.array_equal(ev, g[f"out_events_{k}"]), k


def test_cascade_catalogue_is_seeded(kamd):
    a = kamd.cascade(3000, 2, 3, 0.1)
    b = kamd.cascade(3000, 2, 3, 0.1)
    c = kamd.cascade(3000, 3, 3, 0.1)
    assert len(a) > 3000 and a.tobytes() == b.tobytes() and a.tobytes() != c.tobytes()
