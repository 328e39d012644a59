"""CPU tier: catalogue text I/O of the host library (khost_write_catalog,
khost_load_catalog) against the oracle's loader, which the golden fixtures
pin to the reference's load_catalog + read_epoch_jd (orbital.cpp:150-184).

The writer emits the reference's pipe-separated initial-state format with
17 significant digits, so a generated catalogue loads back to the same
doubles; the loader parses lines in parallel and must agree with the
oracle's sequential parse bit for bit (ids, states, sizes, types, epoch),
on the real catalogue and on a written 1M-line file.
"""
import time

import numpy as np

EPOCH_JD = 2461253.916666667


def _check_against_oracle(kamd, orc, path):
    t0 = time.perf_counter()
    got, ep, dropped = kamd.load_catalog(path)
    t_engine = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref = orc.load_catalog(path)
    t_oracle = time.perf_counter() - t0
    assert len(got) == len(ref["states"]) and dropped == ref["dropped"]
    assert ep == ref["epoch_jd"]
    assert np.array_equal(got["id"], ref["ids"])
    assert np.array_equal(got["state"], ref["states"])
    assert np.array_equal(got["size_m"], ref["size_m"])
    assert np.array_equal(got["type"], ref["type"])
    return got, ep, t_engine, t_oracle


def test_real_catalogue_loads_as_the_oracle_does(kamd, orc, catalog_path):
    got, ep, _, _ = _check_against_oracle(kamd, orc, catalog_path)
    assert len(got) > 20000 and abs(ep - EPOCH_JD) < 1e-9


def test_written_catalogue_round_trips_exactly(kamd, orc, tmp_path):
    objs = kamd.generate(5000, seed=3)
    path = str(tmp_path / "generated.txt")
    kamd.write_catalog(path, objs)
    got, ep, _, _ = _check_against_oracle(kamd, orc, path)
    assert abs(ep - EPOCH_JD) < 1e-9
    assert np.array_equal(got["state"], objs["state"])
    assert np.array_equal(got["size_m"], objs["size_m"])


def test_million_line_catalogue(kamd, orc, tmp_path):
    objs = kamd.generate(1_000_000, seed=0, mix=(0.7, 0.1, 0.1, 0.1, 1.5))
    path = str(tmp_path / "million.txt")
    t0 = time.perf_counter()
    kamd.write_catalog(path, objs)
    t_write = time.perf_counter() - t0
    got, _, t_engine, t_oracle = _check_against_oracle(kamd, orc, path)
    print(f"1M lines: write {t_write:.2f} s, parallel load {t_engine:.2f} s, oracle load {t_oracle:.2f} s")
    assert np.array_equal(got["state"], objs["state"])
