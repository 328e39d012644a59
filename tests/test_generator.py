"""CPU tier: the seeded synthetic catalogue (khost_generate) that defines every
benchmark workload, pinned against the orbit mix it is specified to draw.

Elements are recovered from the generated ECI states by textbook two-body
formulas written here (not by the generator's code), and each stated
distribution is checked with a Kolmogorov-Smirnov test against its exact CDF
or with binomial bounds on class fractions:

* orbit classes 70 / 10 / 10 / 10 % (LEO / GTO-HEO / MEO / GEO);
* LEO semi-major-axis altitude: 45 % N(550, 15), 30 % N(800, 40), 25 %
  log-uniform 300-2000 km, clamped to 300-2000 km; inclination 40 %
  N(53, 0.3) deg, 30 % U(97, 99), 30 % U(0, 100); eccentricity Exp(mean
  0.005) (perigee kept above 200 km);
* GTO/HEO perigee altitude U(200, 600) km, e U(0.6, 0.75), inclination 60 %
  U(0, 28.5) deg, 40 % N(63.4, 0.5);
* MEO a - R_E N(20200, 30) km at N(55, 0.5) deg; GEO a - R_E N(35786, 20)
  km, inclination U(0, 15) deg;
* RAAN, argument of perigee and mean anomaly uniform on [0, 2 pi);
* size classes 0.2 / 0.8 / 4 m at 50 / 30 / 20 %.

The Kepler -> ECI conversion itself (khost_kepler_state) is pinned by a
round trip through the independent element formulas and by the two-body
invariants (energy -mu / 2a, angular momentum sqrt(mu a (1 - e^2)) along
(sin i sin RAAN, -sin i cos RAAN, cos i)).
"""
import numpy as np
import pytest
from scipy import stats

MU = 398600.4418
R_E = 6378.137
N = 200_000
DEG = np.pi / 180.0


def elements(s):
    """(a, e, i, RAAN, argp, M) of ECI states, angles in radians."""
    r, v = s[:, :3], s[:, 3:]
    rn = np.linalg.norm(r, axis=1)
    en = 0.5 * (v * v).sum(1) - MU / rn
    a = -MU / (2.0 * en)
    h = np.cross(r, v)
    hn = np.linalg.norm(h, axis=1)
    inc = np.arccos(np.clip(h[:, 2] / hn, -1.0, 1.0))
    node = np.stack([-h[:, 1], h[:, 0], np.zeros(len(h))], 1)
    nn = np.linalg.norm(node, axis=1)
    raan = np.mod(np.arctan2(node[:, 1], node[:, 0]), 2 * np.pi)
    evec = np.cross(v, h) / MU - r / rn[:, None]
    e = np.linalg.norm(evec, axis=1)
    argp = np.arccos(np.clip((node * evec).sum(1) / (nn * e), -1.0, 1.0))
    argp = np.where(evec[:, 2] < 0, 2 * np.pi - argp, argp)
    nu = np.arccos(np.clip((evec * r).sum(1) / (e * rn), -1.0, 1.0))
    nu = np.where((r * v).sum(1) < 0, 2 * np.pi - nu, nu)
    E = 2.0 * np.arctan2(np.sqrt(1 - e) * np.sin(nu / 2), np.sqrt(1 + e) * np.cos(nu / 2))
    M = np.mod(E - e * np.sin(E), 2 * np.pi)
    return a, e, inc, raan, argp, M


@pytest.fixture(scope="module")
def cat(kamd):
    pop = kamd.generate(N, seed=0)
    a, e, inc, raan, argp, M = elements(pop["state"])
    cls = np.full(N, -1)
    cls[(a < 12000.0) & (e < 0.3)] = 0                     # LEO
    cls[e > 0.5] = 1                                       # GTO / HEO
    cls[(np.abs(a - R_E - 20200.0) < 400.0) & (e < 0.1)] = 2  # MEO
    cls[(np.abs(a - R_E - 35786.0) < 400.0) & (e < 0.1)] = 3  # GEO
    return dict(pop=pop, a=a, e=e, inc=inc, raan=raan, argp=argp, M=M, cls=cls)


def binomial_ok(count, n, p, sigmas=4.5):
    return abs(count - n * p) <= sigmas * np.sqrt(n * p * (1 - p))


def ks_ok(x, cdf, alpha=1e-4):
    d, pval = stats.kstest(x, cdf)
    return pval > alpha, d, pval


def test_class_fractions_and_sizes(cat):
    cls = cat["cls"]
    assert (cls >= 0).all(), "every object falls in one orbit class"
    for k, p in enumerate((0.7, 0.1, 0.1, 0.1)):
        assert binomial_ok(int((cls == k).sum()), N, p), (k, int((cls == k).sum()))
    size = cat["pop"]["size_m"]
    assert set(np.unique(size)) == {0.2, 0.8, 4.0}
    for s, p in ((0.2, 0.5), (0.8, 0.3), (4.0, 0.2)):
        assert binomial_ok(int((size == s).sum()), N, p), s
    # Type follows the size class (LARGE -> payload).
    assert ((cat["pop"]["type"] == 0) == (size == 4.0)).all()


def test_leo_altitude_mixture(cat):
    alt = cat["a"][cat["cls"] == 0] - R_E
    assert alt.min() >= 300.0 - 1e-6 and alt.max() <= 2000.0 + 1e-6
    w = np.array([0.45, 0.30, 0.25])

    def cdf(x):
        x = np.asarray(x, dtype=np.float64)
        lu = np.clip(np.log(np.maximum(x, 1e-9) / 300.0) / np.log(2000.0 / 300.0), 0.0, 1.0)
        f = w[0] * stats.norm.cdf(x, 550.0, 15.0) + w[1] * stats.norm.cdf(x, 800.0, 40.0) + w[2] * lu
        # Clamping puts the normals' tails on the 300 / 2000 km edges.
        return np.where(x >= 2000.0, 1.0, np.where(x < 300.0, 0.0, f))

    ok, d, p = ks_ok(alt, cdf)
    print(f"LEO altitude KS d={d:.5f} p={p:.3g}")
    assert ok
    # The two shells are the histogram's peaks.
    h, edges = np.histogram(alt, bins=np.arange(300.0, 2001.0, 10.0))
    centres = 0.5 * (edges[1:] + edges[:-1])
    assert abs(centres[np.argmax(h)] - 550.0) <= 10.0
    far = centres > 650.0
    assert abs(centres[far][np.argmax(h[far])] - 800.0) <= 30.0


def test_leo_inclination_and_eccentricity(cat):
    leo = cat["cls"] == 0
    inc = cat["inc"][leo] / DEG

    def cdf(x):
        return (0.4 * stats.norm.cdf(x, 53.0, 0.3) + 0.3 * stats.uniform.cdf(x, 97.0, 2.0) +
                0.3 * stats.uniform.cdf(x, 0.0, 100.0))

    ok, d, p = ks_ok(inc, cdf)
    print(f"LEO inclination KS d={d:.5f} p={p:.3g}")
    assert ok
    # Eccentricity: Exp(mean 0.005) where the perigee cap cannot bind
    # (altitude above 1000 km: the cap is then above 0.1).
    high = leo & (cat["a"] - R_E > 1000.0)
    ok, d, p = ks_ok(cat["e"][high], stats.expon(scale=0.005).cdf)
    print(f"LEO eccentricity KS d={d:.5f} p={p:.3g}, mean {cat['e'][leo].mean():.5f}")
    assert ok
    perigee = cat["a"][leo] * (1 - cat["e"][leo]) - R_E
    assert perigee.min() >= 200.0 - 1e-6


def test_heo_meo_geo(cat):
    heo = cat["cls"] == 1
    hp = cat["a"][heo] * (1 - cat["e"][heo]) - R_E
    for x, dist, name in ((hp, stats.uniform(200.0, 400.0), "GTO perigee altitude"),
                          (cat["e"][heo], stats.uniform(0.6, 0.15), "GTO eccentricity")):
        ok, d, p = ks_ok(x, dist.cdf)
        print(f"{name} KS d={d:.5f} p={p:.3g}")
        assert ok, name
    inc_h = cat["inc"][heo] / DEG
    ok, d, p = ks_ok(inc_h, lambda x: 0.6 * stats.uniform.cdf(x, 0.0, 28.5) + 0.4 * stats.norm.cdf(x, 63.4, 0.5))
    assert ok, "GTO inclination"
    meo = cat["cls"] == 2
    ok, _, _ = ks_ok(cat["a"][meo] - R_E, stats.norm(20200.0, 30.0).cdf)
    assert ok, "MEO radius"
    ok, _, _ = ks_ok(cat["inc"][meo] / DEG, stats.norm(55.0, 0.5).cdf)
    assert ok, "MEO inclination"
    geo = cat["cls"] == 3
    ok, _, _ = ks_ok(cat["a"][geo] - R_E, stats.norm(35786.0, 20.0).cdf)
    assert ok, "GEO radius"
    ok, _, _ = ks_ok(cat["inc"][geo] / DEG, stats.uniform(0.0, 15.0).cdf)
    assert ok, "GEO inclination"


def test_angles_uniform(cat):
    # Well-defined angles: GTO/HEO (large e) for the argument of perigee and
    # mean anomaly; every class above 1 deg inclination for the RAAN.
    heo = cat["cls"] == 1
    tilted = cat["inc"] > 1.0 * DEG
    u = stats.uniform(0.0, 2 * np.pi).cdf
    for x, name in ((cat["raan"][tilted], "RAAN"), (cat["argp"][heo], "argument of perigee"),
                    (cat["M"][heo], "mean anomaly")):
        ok, d, p = ks_ok(x, u)
        print(f"{name} KS d={d:.5f} p={p:.3g}")
        assert ok, name


def test_kepler_round_trip_and_invariants(kamd):
    rng = np.random.default_rng(3)
    n = 4000
    el = np.stack([rng.uniform(6700.0, 45000.0, n), rng.uniform(0.0, 0.8, n), rng.uniform(0.05, 3.09, n),
                   rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)], 1)
    s = kamd.kepler_state(el)
    a, e, inc, raan, argp, M = elements(s)

    def ang(x, y):
        return np.abs(np.angle(np.exp(1j * (x - y))))

    assert np.allclose(a, el[:, 0], rtol=1e-10)
    assert np.allclose(e, el[:, 1], atol=1e-10)
    assert np.allclose(inc, el[:, 2], atol=1e-10)
    assert ang(raan, el[:, 3]).max() < 1e-9
    big_e = el[:, 1] > 1e-3
    assert ang(argp[big_e], el[big_e, 4]).max() < 1e-6
    assert ang(M[big_e], el[big_e, 5]).max() < 1e-6
    # Two-body invariants, independently of the angle formulas.
    r, v = s[:, :3], s[:, 3:]
    energy = 0.5 * (v * v).sum(1) - MU / np.linalg.norm(r, axis=1)
    assert np.allclose(energy, -MU / (2 * el[:, 0]), rtol=1e-11)
    h = np.cross(r, v)
    hn = np.sqrt(MU * el[:, 0] * (1 - el[:, 1] ** 2))
    want = np.stack([np.sin(el[:, 2]) * np.sin(el[:, 3]), -np.sin(el[:, 2]) * np.cos(el[:, 3]),
                     np.cos(el[:, 2])], 1) * hn[:, None]
    assert np.allclose(h, want, rtol=1e-10, atol=1e-8 * hn.max())
    # Perigee on the argument-of-perigee direction: at M = 0 the position is
    # a (1 - e) along (cos w cos O - sin w sin O cos i, ...).
    el0 = el.copy()
    el0[:, 5] = 0.0
    s0 = kamd.kepler_state(el0)
    w, O, i = el[:, 4], el[:, 3], el[:, 2]
    dirp = np.stack([np.cos(w) * np.cos(O) - np.sin(w) * np.sin(O) * np.cos(i),
                     np.cos(w) * np.sin(O) + np.sin(w) * np.cos(O) * np.cos(i), np.sin(w) * np.sin(i)], 1)
    assert np.allclose(s0[:, :3], dirp * (el[:, 0] * (1 - el[:, 1]))[:, None], rtol=1e-10, atol=1e-6)
