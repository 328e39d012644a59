"""CPU tier: the C-ABI library loads and exports every entry point that
include/kessler_hip.h declares, and fails cleanly where no GPU is present
(no compute call is made here)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(khip_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree(kamd):
    assert sorted(kamd.ABI_SYMBOLS) == _declared("kessler_hip.h")


def test_library_exports_every_declared_symbol(kamd):
    lib = kamd.library()
    for name in _declared("kessler_hip.h"):
        assert hasattr(lib, name), name


def test_create_without_device_fails_cleanly(kamd):
    if kamd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        kamd.Engine(0)
