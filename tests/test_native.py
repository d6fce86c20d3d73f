"""CPU checks of the C ABI: the library builds, loads and exports every declared entry point."""

import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "corrct_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(chip_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "chip_fp" in syms and "chip_bp" in syms and "chip_att_maps" in syms
    assert len(syms) >= 20


def test_library_exports_all_declared_symbols():
    from pycorrectedemissionct_amd import _native

    lib = _native.lib()
    exported = set(_native.exported_symbols())
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert name in exported, name


def test_error_path_without_gpu_work():
    """An invalid plan request fails with a status and a message (no device work is launched)."""
    import ctypes

    from pycorrectedemissionct_amd import _native

    lib = _native.lib()
    h = ctypes.c_void_p()
    st = lib.chip_plan_create(ctypes.byref(h), 1, 4, None, 1, None, None, None)
    assert st != 0
    assert "slice side" in lib.chip_last_error().decode()
    with pytest.raises(_native.NativeError):
        _native.check(st, "chip_plan_create")
