"""Python handle on the MI355X propagate-and-screen engine.

The engine is ``lib/libkessler_hip.so`` (HIP kernels for gfx950 + the C ABI of
``include/kessler_hip.h``); this module only binds it with ctypes for the
tests and the benchmark. There is no Python or CPU fallback: if the library
is missing or no GPU is visible, the calls raise.

The directory name is not a Python identifier; load it with
``__graft_entry__.load_package()`` (importlib, module name ``kessler_amd``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libkessler_hip.so")

EVENT_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("step", "<i8"), ("t", "<f8"),
                        ("miss_km", "<f8"), ("v_rel_km_s", "<f8"), ("collision", "<i4"),
                        ("pad", "<i4")])
STATS_DTYPE = np.dtype([("candidate_pairs", "<i8"), ("refined_pairs", "<i8"),
                        ("co_orbiting", "<i8"), ("cell_km", "<f8"),
                        ("max_cell_objects", "<i4"), ("cell_doublings", "<i4"),
                        ("propagated", "<i8"), ("dense_candidates", "<i8"),
                        ("survivors", "<i8")])
OBJECT_DTYPE = np.dtype([("id", "<i4"), ("event", "<i4"), ("alive", "<i4"), ("type", "<i4"),
                         ("state", "<f8", (6,)), ("size_m", "<f8"), ("mass_kg", "<f8"),
                         ("area_to_mass", "<f8")])
TIMING_DTYPE = np.dtype([("propagate_ms", "<f8"), ("grid_ms", "<f8"), ("pairs_ms", "<f8"),
                         ("pairs_sparse_ms", "<f8"), ("pairs_dense_ms", "<f8"), ("exact_ms", "<f8"),
                         ("total_ms", "<f8"), ("steps", "<i8"), ("reruns", "<i8"),
                         ("rerun_reasons", "<i8"), ("reverify_ms", "<f8")])

# Every symbol include/kessler_hip.h declares.
ABI_SYMBOLS = (
    "khip_device_count", "khip_create", "khip_destroy", "khip_last_error",
    "khip_set_objects", "khip_object_count", "khip_get_objects", "khip_set_config",
    "khip_set_time", "khip_get_time", "khip_run", "khip_screen", "khip_screen_static", "khip_take_events",
    "khip_take_stats", "khip_set_timing", "khip_get_timing", "khip_reset_timing",
    "khip_set_chunk", "khip_fp64_selfcheck", "khip_get_previous", "khip_set_precision",
    "khip_net_train", "khip_net_set", "khip_net_get", "khip_net_infer", "khip_net_gradient",
    "khip_set_learned", "khip_get_learned_report", "khip_set_scenarios", "khip_set_graph",
    "khip_get_graph_counts", "khip_set_tuning", "khip_get_tuning", "khip_autotune",
    "khip_set_breakup", "khip_set_masses", "khip_get_masses", "khip_get_sizes", "khip_take_breakups",
)
BREAKUP_DTYPE = np.dtype([("t", "<f8"), ("v_rel_km_s", "<f8"), ("energy_j_per_g", "<f8"),
                          ("target_mass_kg", "<f8"), ("projectile_mass_kg", "<f8"), ("sbm_mass_kg", "<f8"),
                          ("expected_fragments", "<f8"), ("fragments_mass_kg", "<f8"), ("position", "<f8", (3,)),
                          ("id", "<i4"), ("target", "<i4"), ("projectile", "<i4"), ("catastrophic", "<i4"),
                          ("drawn", "<i4"), ("fragments", "<i4"), ("escaped", "<i4"), ("first_index", "<i4")])
PRECISION_LEARNED = 2
NET_PARAMS = 35206
NET_INPUTS = 10
NET_OUTPUTS = 6
PRECISION_FP64 = 0
PRECISION_MIXED = 1
# Every symbol include/kessler_host.h declares.
HOST_SYMBOLS = (
    "khost_generate", "khost_impacts", "khost_cascade", "khost_write_catalog",
    "khost_load_catalog", "khost_moon_position", "khost_write_snapshot", "khost_read_snapshot",
    "khost_kepler_state",
)

_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double


class NetNorm(C.Structure):
    """khip_net_norm: input / output standardisation and the trained step."""
    _fields_ = [("fmu", C.c_float * 10), ("fsig", C.c_float * 10), ("omu", C.c_float * 6),
                ("osig", C.c_float * 6), ("tau", _D)]


class TrainConfig(C.Structure):
    _fields_ = [("samples_per_orbit", C.c_int32), ("spacing_steps", C.c_int32),
                ("iterations", C.c_int32), ("batch", C.c_int32), ("lr_max", _D), ("e_max", _D),
                ("a_max_km", _D), ("tau", _D), ("seed", C.c_uint64)]


class LearnedConfig(C.Structure):
    _fields_ = [("substeps", C.c_int32), ("check_every", C.c_int32), ("tol_km", _D), ("margin_km", _D),
                ("e_max", _D), ("a_max_km", _D), ("tol_a_km", _D), ("windows", C.c_int32), ("margin_factor", _D),
                ("shell_km", _D)]


LEARNED_REPORT_DTYPE = np.dtype([("learned", "<i8"), ("base", "<i8"), ("plain", "<i8"), ("rejected", "<i8"),
                                 ("dev_max_km", "<f8"), ("checks", "<i8"), ("steps", "<i8"),
                                 ("events_verified", "<i8"), ("events_dropped", "<i8"),
                                 ("events_unverified", "<i8"), ("sample_dev_max_km", "<f8"),
                                 ("sample_dev_v_max_km_s", "<f8"), ("band_km", "<f8"), ("band_v_km_s", "<f8"),
                                 ("sample_objects", "<i8")])


class TrainReport(C.Structure):
    _fields_ = [("samples", C.c_int64), ("loss_first", _D), ("loss_last", _D), ("gen_ms", _D),
                ("train_ms", _D)]


class BreakupConfig(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("reserve", C.c_int32), ("lc_m", _D), ("catastrophic_j_per_g", _D),
                ("seed", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("cell_mult", _D), ("dense_min", C.c_int32), ("long_range", C.c_int32), ("seg", C.c_int32),
                ("dense_blocks", C.c_int32), ("exact_blocks", C.c_int32), ("sparse_kernel", C.c_int32)]


class ScreenConfig(C.Structure):
    _fields_ = [("threshold_km", _D), ("radius_scale", _D), ("min_encounter_speed_km_s", _D)]


def build(jobs: int = 8) -> str:
    """Compiles the engine for gfx950 (always recompiles) and returns the .so path."""
    subprocess.run(["make", "-B", "-C", HERE, f"-j{jobs}"], check=True)
    return LIB_PATH


_lib = None


def library() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP engine not built: {LIB_PATH} is missing (run build())")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "khip_device_count": (_I, []),
        "khip_create": (_I, [_I, C.POINTER(_P), C.c_char_p, C.c_size_t]),
        "khip_destroy": (None, [_P]),
        "khip_last_error": (C.c_char_p, [_P]),
        "khip_set_objects": (_I, [_P, _L, _P, _P, _P]),
        "khip_object_count": (_L, [_P]),
        "khip_get_objects": (_I, [_P, _P, _P]),
        "khip_set_config": (_I, [_P, C.POINTER(ScreenConfig)]),
        "khip_set_time": (_I, [_P, _D, _D, _D]),
        "khip_get_time": (_I, [_P, C.POINTER(_D), C.POINTER(_L)]),
        "khip_run": (_I, [_P, _L, _I]),
        "khip_screen": (_I, [_P, _L, _P, _P, _P, _P, _D, _D, _P, _L, C.POINTER(_L), _P]),
        "khip_screen_static": (_I, [_P, _L, _P, _P, _P, _P, _D, _D, _P, _L, C.POINTER(_L), _P]),
        "khip_take_events": (_I, [_P, _P, _L, C.POINTER(_L)]),
        "khip_take_stats": (_I, [_P, _P, _L, C.POINTER(_L)]),
        "khip_set_timing": (_I, [_P, _I]),
        "khip_get_timing": (_I, [_P, _P]),
        "khip_reset_timing": (_I, [_P]),
        "khip_set_chunk": (_I, [_P, _I]),
        "khip_fp64_selfcheck": (_I, [_P, _L, _P, _P, _P]),
        "khip_get_previous": (_I, [_P, _L, _P, _P]),
        "khip_set_precision": (_I, [_P, _I]),
        "khip_net_train": (_I, [_P, _L, _P, C.POINTER(TrainConfig), C.POINTER(TrainReport), _P]),
        "khip_net_set": (_I, [_P, _P, C.POINTER(NetNorm)]),
        "khip_net_get": (_I, [_P, _P, C.POINTER(NetNorm)]),
        "khip_net_infer": (_I, [_P, _L, _P, _P]),
        "khip_net_gradient": (_I, [_P, _L, _P, _P, _P, _P]),
        "khip_set_learned": (_I, [_P, C.POINTER(LearnedConfig)]),
        "khip_get_learned_report": (_I, [_P, _P]),
        "khip_set_scenarios": (_I, [_P, C.c_int32, _P]),
        "khip_set_graph": (_I, [_P, _I]),
        "khip_get_graph_counts": (_I, [_P, C.POINTER(_L), C.POINTER(_L)]),
        "khip_set_tuning": (_I, [_P, C.POINTER(Tuning)]),
        "khip_set_breakup": (_I, [_P, C.POINTER(BreakupConfig)]),
        "khip_set_masses": (_I, [_P, _L, _P]),
        "khip_get_masses": (_I, [_P, _L, _P]),
        "khip_get_sizes": (_I, [_P, _L, _P]),
        "khip_take_breakups": (_I, [_P, _P, _L, C.POINTER(_L)]),
        "khip_get_tuning": (_I, [_P, C.POINTER(Tuning)]),
        "khip_autotune": (_I, [_P, _L, _I, C.POINTER(Tuning), C.POINTER(_D), C.POINTER(_D)]),
        "khost_generate": (_I, [_L, C.c_uint64, _P, _P]),
        "khost_impacts": (_L, [_L, _P, _D, _D, _I, _P, _P, _P, C.c_uint64, _D, _L, _P, _P]),
        "khost_cascade": (_L, [_L, C.c_uint64, _I, _D, _D, _D, _D, _D, _L, _P]),
        "khost_write_catalog": (_I, [C.c_char_p, _L, _P, C.c_char_p]),
        "khost_load_catalog": (_L, [C.c_char_p, _L, _P, C.POINTER(_D), C.POINTER(_I)]),
        "khost_moon_position": (None, [_D, _P]),
        "khost_kepler_state": (None, [_P, _P]),
        "khost_write_snapshot": (_I, [C.c_char_p, _L, _P, _D, _I]),
        "khost_read_snapshot": (_L, [C.c_char_p, _L, _L, _P, C.POINTER(_D)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class Engine:
    """One device context (khip_ctx). Mirrors the reference's Simulation +
    Screener pair for one population: upload once, run steps, read back."""

    def __init__(self, device: int = 0):
        self.lib = library()
        msg = C.create_string_buffer(512)
        h = _P()
        rc = self.lib.khip_create(device, C.byref(h), msg, 512)
        if rc != 0:
            raise RuntimeError(f"khip_create failed ({rc}): {msg.value.decode()}")
        self.h = h
        self.device_name = msg.value.decode()

    def close(self):
        if getattr(self, "h", None):
            self.lib.khip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            err = self.lib.khip_last_error(self.h)
            raise RuntimeError(f"{what} failed ({rc}): {err.decode() if err else ''}")

    def set_objects(self, states, alive=None, size_m=None):
        states = np.ascontiguousarray(states, dtype=np.float64)
        n = states.shape[0]
        alive = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        size_m = None if size_m is None else np.ascontiguousarray(size_m, dtype=np.float64)
        self._check(self.lib.khip_set_objects(self.h, n, _ptr(states), _ptr(alive), _ptr(size_m)),
                    "khip_set_objects")

    @property
    def n(self) -> int:
        return int(self.lib.khip_object_count(self.h))

    def get_objects(self):
        n = self.n
        states = np.zeros((n, 6))
        alive = np.zeros(n, dtype=np.uint8)
        self._check(self.lib.khip_get_objects(self.h, _ptr(states), _ptr(alive)), "khip_get_objects")
        return states, alive

    def set_config(self, threshold_km=5.0, radius_scale=1.0, min_encounter_speed_km_s=0.01):
        cfg = ScreenConfig(threshold_km, radius_scale, min_encounter_speed_km_s)
        self._check(self.lib.khip_set_config(self.h, C.byref(cfg)), "khip_set_config")

    def set_time(self, epoch_jd, t=0.0, dt=10.0):
        self._check(self.lib.khip_set_time(self.h, float(epoch_jd), float(t), float(dt)),
                    "khip_set_time")

    def time(self):
        t = _D()
        s = _L()
        self._check(self.lib.khip_get_time(self.h, C.byref(t), C.byref(s)), "khip_get_time")
        return t.value, s.value

    def set_precision(self, mode: int):
        self._check(self.lib.khip_set_precision(self.h, int(mode)), "khip_set_precision")

    def set_chunk(self, steps: int):
        self._check(self.lib.khip_set_chunk(self.h, int(steps)), "khip_set_chunk")

    def run(self, nsteps: int, screen: bool = True):
        self._check(self.lib.khip_run(self.h, int(nsteps), 1 if screen else 0), "khip_run")

    def take_events(self) -> np.ndarray:
        n = _L()
        self._check(self.lib.khip_take_events(self.h, None, 0, C.byref(n)), "khip_take_events")
        out = np.zeros(n.value, dtype=EVENT_DTYPE)
        self._check(self.lib.khip_take_events(self.h, _ptr(out), n.value, C.byref(n)),
                    "khip_take_events")
        return out[: n.value]

    def take_stats(self) -> np.ndarray:
        n = _L()
        self._check(self.lib.khip_take_stats(self.h, None, 0, C.byref(n)), "khip_take_stats")
        out = np.zeros(n.value, dtype=STATS_DTYPE)
        self._check(self.lib.khip_take_stats(self.h, _ptr(out), n.value, C.byref(n)),
                    "khip_take_stats")
        return out[: n.value]

    def screen(self, before, after, alive_after, size_m, t0, dt, cap=1 << 20):
        before = np.ascontiguousarray(before, dtype=np.float64)
        after = np.ascontiguousarray(after, dtype=np.float64)
        alive = np.ascontiguousarray(alive_after, dtype=np.uint8)
        size_m = np.ascontiguousarray(size_m, dtype=np.float64)
        out = np.zeros(cap, dtype=EVENT_DTYPE)
        st = np.zeros(1, dtype=STATS_DTYPE)
        ne = _L()
        self._check(self.lib.khip_screen(self.h, before.shape[0], _ptr(before), _ptr(after),
                                         _ptr(alive), _ptr(size_m), float(t0), float(dt),
                                         _ptr(out), cap, C.byref(ne), _ptr(st)), "khip_screen")
        assert ne.value <= cap, "event buffer too small"
        return out[: ne.value].copy(), st[0]

    def get_previous(self, idx):
        """Start-of-step states of the last step run, for objects `idx`."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((len(idx), 6))
        self._check(self.lib.khip_get_previous(self.h, len(idx), _ptr(idx), _ptr(out)),
                    "khip_get_previous")
        return out

    def set_timing(self, on: bool):
        self._check(self.lib.khip_set_timing(self.h, 1 if on else 0), "khip_set_timing")

    def timing(self):
        out = np.zeros(1, dtype=TIMING_DTYPE)
        self._check(self.lib.khip_get_timing(self.h, _ptr(out)), "khip_get_timing")
        return out[0]

    def reset_timing(self):
        self._check(self.lib.khip_reset_timing(self.h), "khip_reset_timing")

    # ---- learned corrector -------------------------------------------------

    def net_train(self, orbits, samples_per_orbit=10, spacing_steps=12, iterations=20000,
                  batch=16384, lr_max=3e-3, e_max=0.1, a_max_km=12000.0, tau=60.0, seed=1):
        """Generates the training set from `orbits` (n x 6) on the device and
        trains the corrector there. Returns (report dict, loss curve)."""
        orbits = np.ascontiguousarray(orbits, dtype=np.float64)
        cfg = TrainConfig(samples_per_orbit, spacing_steps, iterations, batch, lr_max, e_max, a_max_km, tau,
                          seed)
        rep = TrainReport()
        curve = np.zeros(iterations, dtype=np.float32)
        self._check(self.lib.khip_net_train(self.h, orbits.shape[0], _ptr(orbits), C.byref(cfg),
                                            C.byref(rep), _ptr(curve)), "khip_net_train")
        return {k: getattr(rep, k) for k, _ in TrainReport._fields_}, curve

    def net_set(self, params, norm: dict):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.shape == (NET_PARAMS,)
        nn = NetNorm()
        for k in ("fmu", "fsig", "omu", "osig"):
            getattr(nn, k)[:] = [float(v) for v in norm[k]]
        nn.tau = float(norm["tau"])
        self._check(self.lib.khip_net_set(self.h, _ptr(params), C.byref(nn)), "khip_net_set")

    def net_get(self):
        params = np.zeros(NET_PARAMS, dtype=np.float32)
        nn = NetNorm()
        self._check(self.lib.khip_net_get(self.h, _ptr(params), C.byref(nn)), "khip_net_get")
        norm = {k: np.array(getattr(nn, k)[:], dtype=np.float32) for k in ("fmu", "fsig", "omu", "osig")}
        norm["tau"] = nn.tau
        return params, norm

    def net_infer(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], NET_OUTPUTS), dtype=np.float32)
        self._check(self.lib.khip_net_infer(self.h, x.shape[0], _ptr(x), _ptr(out)), "khip_net_infer")
        return out

    def net_gradient(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        grad = np.zeros(NET_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.khip_net_gradient(self.h, x.shape[0], _ptr(x), _ptr(y), _ptr(grad),
                                               C.byref(loss)), "khip_net_gradient")
        return grad, loss.value

    def set_learned(self, substeps=6, check_every=10, tol_km=1e-5, margin_km=0.025, e_max=0.1,
                    a_max_km=12000.0, tol_a_km=5e-8, windows=0, shell_km=16.0, margin_factor=4.0):
        cfg = LearnedConfig(substeps, check_every, tol_km, margin_km, e_max, a_max_km, tol_a_km, windows,
                            margin_factor, shell_km)
        self._check(self.lib.khip_set_learned(self.h, C.byref(cfg)), "khip_set_learned")

    def set_scenarios(self, offsets):
        """Objects [offsets[k], offsets[k+1]) form scenario k (screened separately)."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self._check(self.lib.khip_set_scenarios(self.h, len(off) - 1, _ptr(off)), "khip_set_scenarios")

    def set_graph(self, on: bool):
        self._check(self.lib.khip_set_graph(self.h, 1 if on else 0), "khip_set_graph")

    def graph_counts(self):
        c, l = _L(), _L()
        self._check(self.lib.khip_get_graph_counts(self.h, C.byref(c), C.byref(l)), "khip_get_graph_counts")
        return c.value, l.value

    # ---- in-run breakups -------------------------------------------------------

    def set_breakup(self, enabled=True, reserve=65536, lc_m=0.1, catastrophic_j_per_g=40.0, seed=1):
        cfg = BreakupConfig(1 if enabled else 0, int(reserve), float(lc_m), float(catastrophic_j_per_g), int(seed))
        self._check(self.lib.khip_set_breakup(self.h, C.byref(cfg)), "khip_set_breakup")

    def set_masses(self, mass_kg):
        m = np.ascontiguousarray(mass_kg, dtype=np.float64)
        self._check(self.lib.khip_set_masses(self.h, len(m), _ptr(m)), "khip_set_masses")

    def masses(self):
        m = np.zeros(self.n)
        self._check(self.lib.khip_get_masses(self.h, len(m), _ptr(m)), "khip_get_masses")
        return m

    def sizes(self):
        m = np.zeros(self.n)
        self._check(self.lib.khip_get_sizes(self.h, len(m), _ptr(m)), "khip_get_sizes")
        return m

    def take_breakups(self) -> np.ndarray:
        n = _L()
        self._check(self.lib.khip_take_breakups(self.h, None, 0, C.byref(n)), "khip_take_breakups")
        out = np.zeros(n.value, dtype=BREAKUP_DTYPE)
        self._check(self.lib.khip_take_breakups(self.h, _ptr(out), n.value, C.byref(n)), "khip_take_breakups")
        return out[: n.value]

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self.lib.khip_get_tuning(self.h, C.byref(t)), "khip_get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = Tuning(cur["cell_mult"], cur["dense_min"], cur["long_range"], cur["seg"], cur["dense_blocks"],
                   cur["exact_blocks"], cur["sparse_kernel"])
        self._check(self.lib.khip_set_tuning(self.h, C.byref(t)), "khip_set_tuning")

    def autotune(self, steps=4, allow_cell=False):
        """Returns (best tuning dict, base ms, best ms) for `steps` screened steps."""
        t = Tuning()
        b0, b1 = _D(), _D()
        self._check(self.lib.khip_autotune(self.h, int(steps), 1 if allow_cell else 0, C.byref(t), C.byref(b0),
                                           C.byref(b1)), "khip_autotune")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}, b0.value, b1.value

    def learned_report(self):
        out = np.zeros(1, dtype=LEARNED_REPORT_DTYPE)
        self._check(self.lib.khip_get_learned_report(self.h, _ptr(out)), "khip_get_learned_report")
        return {k: out[0][k].item() for k in LEARNED_REPORT_DTYPE.names}

    def fp64_selfcheck(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros((a.shape[0], 3))
        self._check(self.lib.khip_fp64_selfcheck(self.h, a.shape[0], _ptr(a), _ptr(b), _ptr(out)),
                    "khip_fp64_selfcheck")
        return out


def device_count() -> int:
    return int(library().khip_device_count())


class Mix(C.Structure):
    _fields_ = [("frac_leo", _D), ("frac_heo", _D), ("frac_meo", _D), ("frac_geo", _D),
                ("leo_shell_boost", _D)]


def generate(n: int, seed: int = 0, mix=None) -> np.ndarray:
    """Seeded synthetic catalogue (khost_generate) as OBJECT_DTYPE records."""
    out = np.zeros(n, dtype=OBJECT_DTYPE)
    m = None if mix is None else C.byref(Mix(*mix))
    rc = library().khost_generate(int(n), int(seed), m, _ptr(out))
    if rc != 0:
        raise RuntimeError(f"khost_generate failed ({rc})")
    return out


def kepler_state(elements) -> np.ndarray:
    """ECI state of Kepler elements (a, e, i, RAAN, argp, M; radians) as the generator converts them."""
    el = np.ascontiguousarray(elements, dtype=np.float64).reshape(-1, 6)
    out = np.zeros_like(el)
    lib = library()
    for k in range(len(el)):
        lib.khost_kepler_state(_ptr(el[k]), _ptr(out[k]))
    return out


def impacts(objs, epoch_jd, t, targets, projectile_mass_kg, v_rel_km_s, seed, lc_m):
    """Forced impacts in sequence (khost_impacts): returns (objects, events[fragments, escaped])."""
    objs = np.ascontiguousarray(objs, dtype=OBJECT_DTYPE)
    targets = np.ascontiguousarray(targets, dtype=np.int32)
    pm = np.ascontiguousarray(projectile_mass_kg, dtype=np.float64)
    vr = np.ascontiguousarray(v_rel_km_s, dtype=np.float64)
    lib = library()
    args = (len(objs), _ptr(objs), float(epoch_jd), float(t), len(targets), _ptr(targets), _ptr(pm),
            _ptr(vr), int(seed), float(lc_m))
    total = lib.khost_impacts(*args, 0, None, None)
    if total < 0:
        raise RuntimeError(f"khost_impacts failed ({total})")
    out = np.zeros(total, dtype=OBJECT_DTYPE)
    ev = np.zeros((len(targets), 2), dtype=np.int32)
    lib.khost_impacts(*args, total, _ptr(out), _ptr(ev))
    return out, ev


def cascade(n_catalog: int, seed: int, n_events: int, lc_m: float, mass=(100.0, 2000.0),
            vrel=(8.0, 12.0)) -> np.ndarray:
    """Seeded cascade catalogue (khost_cascade): catalogue + forced impacts at t = 0."""
    lib = library()
    args = (int(n_catalog), int(seed), int(n_events), float(lc_m), float(mass[0]), float(mass[1]),
            float(vrel[0]), float(vrel[1]))
    total = lib.khost_cascade(*args, 0, None)
    if total < 0:
        raise RuntimeError(f"khost_cascade failed ({total})")
    out = np.zeros(total, dtype=OBJECT_DTYPE)
    lib.khost_cascade(*args, total, _ptr(out))
    return out


def write_catalog(path: str, objs, epoch_iso: str = "2026-08-01T10:00:00+00:00"):
    objs = np.ascontiguousarray(objs, dtype=OBJECT_DTYPE)
    rc = library().khost_write_catalog(path.encode(), len(objs), _ptr(objs), epoch_iso.encode())
    if rc != 0:
        raise RuntimeError(f"khost_write_catalog failed ({rc})")


def load_catalog(path: str):
    lib = library()
    ep = _D()
    dr = _I()
    n = lib.khost_load_catalog(path.encode(), 0, None, C.byref(ep), C.byref(dr))
    out = np.zeros(max(n, 0), dtype=OBJECT_DTYPE)
    lib.khost_load_catalog(path.encode(), n, _ptr(out), C.byref(ep), C.byref(dr))
    return out, ep.value, dr.value


def write_snapshot(path: str, objs, t: float, exact: bool = False):
    """write_snapshot / write_snapshot_exact of the host core (khost_write_snapshot)."""
    objs = np.ascontiguousarray(objs, dtype=OBJECT_DTYPE)
    rc = library().khost_write_snapshot(path.encode(), len(objs), _ptr(objs), float(t), 1 if exact else 0)
    if rc != 0:
        raise RuntimeError(f"khost_write_snapshot failed ({rc})")


def read_snapshot(path: str, objs):
    """read_snapshot of the host core (khost_read_snapshot): returns (objects
    after the replay, t), or None where the snapshot does not match `objs`."""
    objs = np.array(objs, dtype=OBJECT_DTYPE, copy=True)
    t = _D()
    cap = len(objs)
    while True:
        buf = np.zeros(max(cap, 1), dtype=OBJECT_DTYPE)
        buf[: len(objs)] = objs
        n = library().khost_read_snapshot(path.encode(), len(objs), cap, _ptr(buf), C.byref(t))
        if n < 0:
            return None
        if n <= cap:
            return buf[:n], t.value
        cap = n


def moon_position(jd: float) -> np.ndarray:
    out = np.zeros(3)
    library().khost_moon_position(float(jd), _ptr(out))
    return out
