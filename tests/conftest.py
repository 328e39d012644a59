import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def orc():
    from oracle.bind import oracle

    o = oracle()
    o.threads(min(16, os.cpu_count() or 1))
    return o


@pytest.fixture(scope="session")
def ref():
    from oracle.bind import reference

    r = reference()
    if r is not None:
        r.threads(min(16, os.cpu_count() or 1))
    return r


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(GOLDEN, "ref_vectors.npz"))


@pytest.fixture(scope="session")
def golden_1h():
    return np.load(os.path.join(GOLDEN, "ref_catalog_1h.npz"))


@pytest.fixture(scope="session")
def catalog_path(tmp_path_factory):
    """The reference's real catalogue (reduced extract), as a plain text file."""
    dst = tmp_path_factory.mktemp("catalog") / "satellites_20260801_1000Z.txt"
    with gzip.open(os.path.join(GOLDEN, "catalog_20260801_1000Z.psv.gz"), "rb") as src, \
            open(dst, "wb") as out:
        shutil.copyfileobj(src, out)
    return str(dst)


@pytest.fixture(scope="session")
def catalog(orc, catalog_path):
    return orc.load_catalog(catalog_path)


@pytest.fixture(scope="session")
def kamd():
    from __graft_entry__ import load_package

    return load_package()


@pytest.fixture()
def engine(kamd):
    eng = kamd.Engine(0)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def trained(kamd):
    """The corrector trained on the device as the benchmark trains it (seeded
    orbits of generator seed 101, 20,000 Adam steps): (report, loss curve,
    parameters, standardisation). Shared by every GPU module that needs it."""
    eng = kamd.Engine(0)
    orbits = kamd.generate(500000, seed=101)["state"]
    eng.set_time(2461253.9166666665, 0.0, 60.0)
    rep, curve = eng.net_train(orbits, samples_per_orbit=15, spacing_steps=12, iterations=20000,
                               batch=16384, lr_max=3e-3, seed=1)
    params, norm = eng.net_get()
    eng.close()
    return rep, curve, params, norm
