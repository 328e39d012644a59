"""CPU tier: the Python mirrors of the C ABI's structs match include/kessler_hip.h.

A small C program, compiled here with gcc against the header, prints the
size of every public struct and the offset of each field (in declaration
order); each ctypes structure and numpy record type of the package must have
the same size and the same offsets, field by field. A field added on one side
only, or a change of type or order, fails here instead of corrupting calls.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# struct -> C fields in declaration order (arrays by their first element).
C_FIELDS = {
    "khip_screen_config": ["threshold_km", "radius_scale", "min_encounter_speed_km_s"],
    "khip_event": ["i", "j", "step", "t", "miss_km", "v_rel_km_s", "collision", "pad"],
    "khip_stats": ["candidate_pairs", "refined_pairs", "co_orbiting", "cell_km", "max_cell_objects",
                   "cell_doublings", "propagated", "dense_candidates", "survivors"],
    "khip_timing": ["propagate_ms", "grid_ms", "pairs_ms", "pairs_sparse_ms", "pairs_dense_ms", "exact_ms",
                    "total_ms", "steps", "reruns", "rerun_reasons", "reverify_ms"],
    "khip_learned_config": ["substeps", "check_every", "tol_km", "margin_km", "e_max", "a_max_km", "tol_a_km",
                            "windows", "margin_factor", "shell_km"],
    "khip_learned_report": ["learned", "base", "plain", "rejected", "dev_max_km", "checks", "steps",
                            "events_verified", "events_dropped", "events_unverified", "sample_dev_max_km",
                            "sample_dev_v_max_km_s", "band_km", "band_v_km_s", "sample_objects"],
    "khip_breakup_config": ["enabled", "reserve", "lc_m", "catastrophic_j_per_g", "seed"],
    "khip_breakup_event": ["t", "v_rel_km_s", "energy_j_per_g", "target_mass_kg", "projectile_mass_kg",
                           "sbm_mass_kg", "expected_fragments", "fragments_mass_kg", "position", "id", "target",
                           "projectile", "catastrophic", "drawn", "fragments", "escaped", "first_index"],
    "khip_tuning": ["cell_mult", "dense_min", "long_range", "seg", "dense_blocks", "exact_blocks",
                    "sparse_kernel"],
    "khip_net_norm": ["fmu", "fsig", "omu", "osig", "tau"],
    "khip_train_config": ["samples_per_orbit", "spacing_steps", "iterations", "batch", "lr_max", "e_max",
                          "a_max_km", "tau", "seed"],
    "khip_train_report": ["samples", "loss_first", "loss_last", "gen_ms", "train_ms"],
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "kessler_hip.h"', "int main(void) {"]
    for s, fields in C_FIELDS.items():
        lines.append(f'  printf("{s} sizeof %zu\\n", sizeof({s}));')
        for f in fields:
            lines.append(f'  printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.split("\n"):
        if line:
            s, f, v = line.split()
            layout.setdefault(s, {})[f] = int(v)
    return layout


def _mirrors(kamd):
    return {
        "khip_screen_config": kamd.ScreenConfig, "khip_event": kamd.EVENT_DTYPE, "khip_stats": kamd.STATS_DTYPE,
        "khip_timing": kamd.TIMING_DTYPE, "khip_learned_config": kamd.LearnedConfig,
        "khip_learned_report": kamd.LEARNED_REPORT_DTYPE, "khip_breakup_config": kamd.BreakupConfig,
        "khip_breakup_event": kamd.BREAKUP_DTYPE, "khip_tuning": kamd.Tuning, "khip_net_norm": kamd.NetNorm,
        "khip_train_config": kamd.TrainConfig, "khip_train_report": kamd.TrainReport,
    }


def _py_layout(m):
    if isinstance(m, np.dtype):
        return m.itemsize, [m.fields[n][1] for n in m.names]
    import ctypes
    return ctypes.sizeof(m), [getattr(m, n).offset for n, _ in m._fields_]


@pytest.mark.parametrize("struct", sorted(C_FIELDS))
def test_python_mirror_matches_header(kamd, c_layout, struct):
    c = c_layout[struct]
    size, offsets = _py_layout(_mirrors(kamd)[struct])
    assert size == c["sizeof"], f"{struct}: sizeof {size} vs {c['sizeof']} in C"
    assert offsets == [c[f] for f in C_FIELDS[struct]], f"{struct}: field offsets differ from the header"
