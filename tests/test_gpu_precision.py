"""GPU tier: the precision form (KHIP_PRECISION_MIXED: FP64 state and central
term, FP32 J2 and Moon) against the FP64 path, which is bit-identical to the
oracle (tests/test_gpu_parity.py).

Bars (1M objects, LEO shells densified, 24 h): 24 h position error <= 10 m; screening at
2 km: recall of the FP64 event set, TCA within 1 ms and miss within 1 m on
the matched events. The oracle anchors the FP64 side here too: a sample of
objects propagated 24 h by the CPU oracle equals the engine's FP64 states.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPOCH = 2461253.9166666665
MIX_DENSE = (0.7, 0.1, 0.1, 0.1, 1.5)  # LEO shells densified


def _run(eng, pop, steps, precision, screen, thr=2.0):
    eng.set_precision(precision)
    eng.set_objects(pop["state"], pop["alive"].astype(np.uint8), pop["size_m"])
    eng.set_config(thr, 1.0, 0.01)
    eng.set_time(EPOCH, 0.0, 10.0)
    eng.run(steps, screen=screen)
    s, a = eng.get_objects()
    return s, a, eng.take_events(), eng.take_stats()


def test_mixed_24h_position_error_under_10m(engine, kamd, orc):
    pop = kamd.generate(20000, seed=5, mix=MIX_DENSE)
    s64, a64, _, _ = _run(engine, pop, 8640, kamd.PRECISION_FP64, screen=False)
    s32, a32, _, _ = _run(engine, pop, 8640, kamd.PRECISION_MIXED, screen=False)
    # FP64 side anchored on the oracle for a sample of objects.
    idx = np.arange(0, 20000, 97)
    so, ao, _, _, _ = orc.run(pop["state"][idx], pop["alive"][idx].astype(np.uint8),
                              pop["size_m"][idx], EPOCH, 10.0, 8640, screen=False, per_step=False)
    assert np.array_equal(so, s64[idx])
    both = (a64 == 1) & (a32 == 1)
    assert np.array_equal(a64, a32)
    dr = np.linalg.norm(s32[both, :3] - s64[both, :3], axis=1) * 1000.0  # m
    dv = np.linalg.norm(s32[both, 3:] - s64[both, 3:], axis=1) * 1e6      # mm/s
    print(f"24 h mixed vs FP64: |dr| max {dr.max():.3f} m p99 {np.percentile(dr, 99):.3f} m; "
          f"|dv| max {dv.max():.3f} mm/s")
    assert dr.max() <= 10.0


def test_mixed_screening_recall_at_2km(engine, kamd):
    pop = kamd.generate(150000, seed=6, mix=MIX_DENSE)
    _, _, e64, t64 = _run(engine, pop, 360, kamd.PRECISION_FP64, screen=True)
    _, _, e32, t32 = _run(engine, pop, 360, kamd.PRECISION_MIXED, screen=True)
    assert len(e64) > 100
    k64 = {(int(a), int(b), int(s)): (t, m) for a, b, s, t, m in
           zip(e64["i"], e64["j"], e64["step"], e64["t"], e64["miss_km"])}
    k32 = {(int(a), int(b), int(s)): (t, m) for a, b, s, t, m in
           zip(e32["i"], e32["j"], e32["step"], e32["t"], e32["miss_km"])}
    common = set(k64) & set(k32)
    recall = len(common) / len(k64)
    dt = max(abs(k64[k][0] - k32[k][0]) for k in common)
    dm = max(abs(k64[k][1] - k32[k][1]) for k in common)
    print(f"1 h at 2 km: {len(k64)} FP64 events, recall {recall:.6f}, extra {len(set(k32) - set(k64))}, "
          f"max |dTCA| {dt * 1e3:.3e} ms, max |dmiss| {dm * 1e3:.3e} m")
    # An event the mixed run misses may only sit on the 2 km threshold, or
    # on a step edge (then the mixed run reports it in the adjacent step).
    for k in set(k64) - set(k32):
        shifted = [(k[0], k[1], k[2] + d) for d in (-1, 1)]
        moved = [x for x in shifted if x in k32 and abs(k32[x][0] - k64[k][0]) < 1e-3]
        assert moved or abs(k64[k][1] - 2.0) < 1e-3, (k, k64[k])
    assert dt < 1e-3 and dm < 1e-3
