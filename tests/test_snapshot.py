"""CPU tier: snapshot files and their replay (write_snapshot,
write_snapshot_exact, read_snapshot; orbital.hpp:125-133) against the
reference's own functions (oracle/_ref, where it was built).

Checked: files written by the host core and by the reference are read the
same way by both readers (object count, ids, states, alive flags, time),
in both directions; the exact variant replays every double; the reference's
matching rules hold (fragments appended past the catalogue, surplus loaded
fragments dropped, a missing catalogue object, a row out of order or a
malformed row rejected).
"""
import numpy as np
import pytest

FRAG = 10_000_000


def _population(kamd, n_cat=300, n_frag=40):
    objs = kamd.generate(n_cat + n_frag, seed=12)
    objs["id"][:n_cat] = np.arange(25544, 25544 + n_cat)
    objs["event"][:n_cat] = -1
    objs["id"][n_cat:] = FRAG + np.arange(n_frag)
    objs["event"][n_cat:] = 0
    objs["alive"] = 1
    objs["alive"][::7] = 0
    return objs


def _both(kamd, ref, path, objs):
    mine = kamd.read_snapshot(path, objs)
    theirs = ref.read_snapshot(path, objs["id"], objs["event"])
    return mine, theirs


def _same(mine, theirs):
    assert (mine is None) == (theirs is None)
    if mine is None:
        return
    got, t = mine
    ids, st, al, t_ref = theirs
    assert len(got) == len(ids) and t == t_ref
    assert np.array_equal(got["id"], ids)
    assert np.array_equal(got["state"], st)
    assert np.array_equal(got["alive"].astype(np.uint8), al)


@pytest.mark.parametrize("writer", ["host", "host_exact", "reference"])
def test_readers_agree_on_every_writer(kamd, ref, tmp_path, writer):
    if ref is None:
        pytest.skip("oracle/_ref is not built")
    objs = _population(kamd)
    path = str(tmp_path / "snap.csv")
    if writer == "reference":
        ref.write_snapshot(path, objs["id"], objs["state"], objs["alive"].astype(np.uint8), 3600.0)
    else:
        kamd.write_snapshot(path, objs, 3600.0, exact=writer == "host_exact")
    # Replay into the full list, into the catalogue only (fragments appended)
    # and into the list plus extra fragments (surplus dropped).
    extra = objs[-5:].copy()
    extra["id"] = FRAG + 1000 + np.arange(5)
    for target in (objs, objs[:300], np.concatenate([objs, extra])):
        mine, theirs = _both(kamd, ref, path, target)
        assert mine is not None
        _same(mine, theirs)
        assert len(mine[0]) == len(objs) and mine[1] == 3600.0
    if writer == "host_exact":
        assert np.array_equal(kamd.read_snapshot(path, objs)[0]["state"], objs["state"])
    else:  # nine significant digits
        got = kamd.read_snapshot(path, objs)[0]["state"]
        assert np.allclose(got, objs["state"], rtol=1e-8, atol=1e-12)


def test_mismatches_are_rejected_as_the_reference_does(kamd, ref, tmp_path):
    if ref is None:
        pytest.skip("oracle/_ref is not built")
    objs = _population(kamd)
    path = str(tmp_path / "snap.csv")
    kamd.write_snapshot(path, objs, 60.0)
    swapped = objs.copy()
    swapped[[3, 4]] = swapped[[4, 3]]              # catalogue order differs
    missing = np.concatenate([objs, objs[:1]])      # a catalogue object past the rows
    missing["id"][-1] = 99999
    missing["event"][-1] = -1
    for target in (swapped, missing):
        mine, theirs = _both(kamd, ref, path, target)
        assert mine is None and theirs is None
    with open(path) as f:
        lines = f.read().splitlines()
    lines[5] = lines[5].rsplit(",", 2)[0]             # a row with too few fields
    bad = str(tmp_path / "bad.csv")
    with open(bad, "w") as f:
        f.write("\n".join(lines) + "\n")
    mine, theirs = _both(kamd, ref, bad, objs)
    assert mine is None and theirs is None
    assert kamd.read_snapshot(str(tmp_path / "absent.csv"), objs) is None
