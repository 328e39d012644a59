"""ctypes bindings of the CPU checkers.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py. Two libraries share one record layout:

* ``liboracle.so``         the C++/OpenMP restatement (oracle/oracle.cpp);
* ``_ref/libkessler_ref.so`` the reference's own core sources compiled as they
  are, plus a forwarding harness (oracle/ref_harness.cpp). Present only when
  it was built in a container that has the reference checkout.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

EVENT_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("step", "<i8"), ("t", "<f8"),
                        ("miss_km", "<f8"), ("v_rel_km_s", "<f8"), ("collision", "<i4"),
                        ("pad", "<i4")])
STATS_DTYPE = np.dtype([("candidate_pairs", "<i8"), ("refined_pairs", "<i8"),
                        ("co_orbiting", "<i8"), ("cell_km", "<f8"),
                        ("max_cell_objects", "<i4"), ("pad", "<i4")])
OBJECT_DTYPE = np.dtype([("id", "<i4"), ("event", "<i4"), ("alive", "<i4"), ("type", "<i4"),
                         ("state", "<f8", (6,)), ("size_m", "<f8"), ("mass_kg", "<f8"),
                         ("area_to_mass", "<f8")])
BREAKUP_DTYPE = np.dtype([("id", "<i4"), ("target", "<i4"), ("projectile", "<i4"),
                          ("catastrophic", "<i4"), ("t", "<f8"), ("v_rel_km_s", "<f8"),
                          ("energy_j_per_g", "<f8"), ("target_mass_kg", "<f8"),
                          ("projectile_mass_kg", "<f8"), ("sbm_mass_kg", "<f8"),
                          ("expected_fragments", "<f8"), ("fragments_mass_kg", "<f8"),
                          ("fragments", "<i4"), ("escaped", "<i4")])

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class Checker:
    """One CPU checker library (prefix ``orc_`` or ``ref_``)."""

    def __init__(self, path: str, prefix: str):
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        f = self._f
        f("configure_threads", _I, [_I])
        f("moon_position", None, [_D, _P])
        f("acceleration", None, [_P, _P, _P])
        f("rk4", None, [_P, _D, _P, _P, _P, _P])
        f("filter_pair", _I, [_P, _P, _D, _D, _D])
        f("prefilter_keep", _I, [_P, _P, C.c_float, C.c_float, C.c_float])
        f("interpolate_state", None, [_P, _P, _D, _D, _P])
        f("screen", _L, [_I, _P, _P, _P, _P, _D, _D, _D, _D, _D, _P, _L, _P])
        f("run", _L, [_I, _P, _P, _P, _D, _D, _D, _L, _I, _D, _D, _D, _P, _L, _P, _P])
        f("load_catalog", _I, [C.c_char_p, _I, _P, _P, _P, _P, _P, _P, _P])

    def _f(self, name, res, args):
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = res
        fn.argtypes = args
        setattr(self, name + "_", fn)

    def threads(self, n: int = 0) -> int:
        return self.configure_threads_(int(n))

    def moon_position(self, jd: float) -> np.ndarray:
        out = np.zeros(3)
        self.moon_position_(float(jd), _ptr(out))
        return out

    def acceleration(self, r, rm) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        rm = np.ascontiguousarray(rm, dtype=np.float64)
        out = np.zeros(3)
        self.acceleration_(_ptr(r), _ptr(rm), _ptr(out))
        return out

    def rk4(self, s, dt, m0, m1, m2) -> np.ndarray:
        s = np.ascontiguousarray(s, dtype=np.float64)
        ms = [np.ascontiguousarray(m, dtype=np.float64) for m in (m0, m1, m2)]
        out = np.zeros(6)
        self.rk4_(_ptr(s), float(dt), _ptr(ms[0]), _ptr(ms[1]), _ptr(ms[2]), _ptr(out))
        return out

    def filter_pair(self, r, v, dt, thr, vmin2) -> int:
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        return self.filter_pair_(_ptr(r), _ptr(v), float(dt), float(thr), float(vmin2))

    def prefilter_keep(self, r, v, dt, thr, vmin2) -> int:
        r = np.ascontiguousarray(r, dtype=np.float32)
        v = np.ascontiguousarray(v, dtype=np.float32)
        return self.prefilter_keep_(_ptr(r), _ptr(v), float(dt), float(thr), float(vmin2))

    def interpolate_state(self, a, b, dt, s) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros(6)
        self.interpolate_state_(_ptr(a), _ptr(b), float(dt), float(s), _ptr(out))
        return out

    def screen(self, before, after, alive_after, size_m, t0, dt, thr=5.0, radius_scale=1.0,
               vmin_km_s=0.01, cap=1 << 20):
        before = np.ascontiguousarray(before, dtype=np.float64)
        after = np.ascontiguousarray(after, dtype=np.float64)
        alive = np.ascontiguousarray(alive_after, dtype=np.uint8)
        size_m = np.ascontiguousarray(size_m, dtype=np.float64)
        n = before.shape[0]
        out = np.zeros(cap, dtype=EVENT_DTYPE)
        st = np.zeros(1, dtype=STATS_DTYPE)
        cnt = self.screen_(n, _ptr(before), _ptr(after), _ptr(alive), _ptr(size_m), float(t0),
                           float(dt), float(thr), float(radius_scale), float(vmin_km_s),
                           _ptr(out), cap, _ptr(st))
        assert cnt <= cap, "event buffer too small"
        return out[:cnt].copy(), st[0]

    def run(self, states, alive, size_m, epoch_jd, dt, nsteps, screen=True, thr=5.0,
            radius_scale=1.0, vmin_km_s=0.01, t_start=0.0, cap=1 << 22, per_step=True):
        """Runs nsteps; returns (states, alive, events, per-step stats, loop_s)."""
        states = np.array(states, dtype=np.float64, order="C", copy=True)
        alive = np.array(alive, dtype=np.uint8, copy=True)
        size_m = np.ascontiguousarray(size_m, dtype=np.float64)
        n = states.shape[0]
        out = np.zeros(cap, dtype=EVENT_DTYPE)
        st = np.zeros(max(int(nsteps), 1), dtype=STATS_DTYPE) if per_step else None
        loop_s = C.c_double(0.0)
        cnt = self.run_(n, _ptr(states), _ptr(alive), _ptr(size_m), float(epoch_jd),
                        float(t_start), float(dt), int(nsteps), 1 if screen else 0, float(thr),
                        float(radius_scale), float(vmin_km_s), _ptr(out), cap, _ptr(st),
                        C.byref(loop_s))
        assert cnt <= cap, "event buffer too small"
        return states, alive, out[:cnt].copy(), st, loop_s.value

    def load_catalog(self, path: str):
        bpath = path.encode()
        n = self.load_catalog_(bpath, 0, None, None, None, None, None, None, None)
        ids = np.zeros(n, dtype=np.int32)
        states = np.zeros((n, 6))
        size_m = np.zeros(n)
        rcs = np.zeros(n, dtype=np.int32)
        typ = np.zeros(n, dtype=np.int32)
        epoch = C.c_double(0.0)
        dropped = C.c_int(0)
        self.load_catalog_(bpath, n, _ptr(ids), _ptr(states), _ptr(size_m), _ptr(rcs), _ptr(typ),
                           C.byref(epoch), C.byref(dropped))
        return dict(ids=ids, states=states, size_m=size_m, rcs=rcs, type=typ,
                    epoch_jd=epoch.value, dropped=dropped.value)


class Reference(Checker):
    """The compiled reference core (oracle/_ref). Adds the breakup entry points."""

    def __init__(self, path: str):
        super().__init__(path, "ref_")
        f = self._f
        f("rng", None, [C.c_uint64, _I, _D, _L, _P])
        f("impacts", _L, [_I, _P, _P, _P, _D, _D, _I, _P, _P, _P, C.c_uint64, _D, _L, _P, _P])
        f("write_snapshot", None, [C.c_char_p, _I, _P, _P, _P, _D])
        f("read_snapshot", _L, [C.c_char_p, _I, _P, _P, _L, _P, _P, C.POINTER(_D)])

    def write_snapshot(self, path, ids, states, alive, t):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        states = np.ascontiguousarray(states, dtype=np.float64)
        alive = np.ascontiguousarray(alive, dtype=np.uint8)
        self.write_snapshot_(path.encode(), len(ids), _ptr(ids), _ptr(states), _ptr(alive), float(t))

    def read_snapshot(self, path, ids, events):
        """(ids, states, alive, t) after the reference's read_snapshot into a
        list with these ids and fragment flags, or None where it returns false."""
        n = len(ids)
        cap = n
        t = _D()
        while True:
            ib = np.zeros(max(cap, 1), dtype=np.int32)
            ib[:n] = ids
            ev = np.ascontiguousarray(events, dtype=np.int32)
            st = np.zeros((max(cap, 1), 6))
            al = np.zeros(max(cap, 1), dtype=np.uint8)
            got = self.read_snapshot_(path.encode(), n, _ptr(ib), _ptr(ev), cap, _ptr(st), _ptr(al), C.byref(t))
            if got < 0:
                return None
            if got <= cap:
                return ib[:got], st[:got], al[:got], t.value
            cap = got

    def rng(self, seed, kind, n, mean=0.0):
        out = np.zeros(n)
        self.rng_(int(seed), int(kind), float(mean), int(n), _ptr(out))
        return out.view(np.uint64) if kind >= 3 else out

    def impacts(self, states, size_m, mass_kg, epoch_jd, t, targets, prj_mass, vrel, seed, lc_m):
        states = np.ascontiguousarray(states, dtype=np.float64)
        size_m = np.ascontiguousarray(size_m, dtype=np.float64)
        mass_kg = np.ascontiguousarray(mass_kg, dtype=np.float64)
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        prj_mass = np.ascontiguousarray(prj_mass, dtype=np.float64)
        vrel = np.ascontiguousarray(vrel, dtype=np.float64)
        n = states.shape[0]
        nimp = targets.shape[0]
        args = (n, _ptr(states), _ptr(size_m), _ptr(mass_kg), float(epoch_jd), float(t), nimp,
                _ptr(targets), _ptr(prj_mass), _ptr(vrel), int(seed), float(lc_m))
        total = self.impacts_(*args, 0, None, None)
        objs = np.zeros(total, dtype=OBJECT_DTYPE)
        evs = np.zeros(nimp, dtype=BREAKUP_DTYPE)
        self.impacts_(*args, total, _ptr(objs), _ptr(evs))
        return objs, evs


class Oracle(Checker):
    """The restatement (liboracle.so). Adds the brute-force pair count."""

    def __init__(self, path: str):
        super().__init__(path, "orc_")
        self._f("bruteforce_candidates", _L, [_I, _P, _P, _D])
        self._f("cell_size", _D, [_I, _P, _P, _D, _D])

    def bruteforce_candidates(self, before, alive, cell) -> int:
        before = np.ascontiguousarray(before, dtype=np.float64)
        alive = np.ascontiguousarray(alive, dtype=np.uint8)
        return int(self.bruteforce_candidates_(before.shape[0], _ptr(before), _ptr(alive), float(cell)))


def oracle() -> Oracle:
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built: run make -C oracle")
    return Oracle(path)


def reference():
    """The compiled reference, or None where it was not built."""
    path = os.path.join(HERE, "_ref", "libkessler_ref.so")
    return Reference(path) if os.path.exists(path) else None
