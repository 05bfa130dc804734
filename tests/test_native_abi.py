"""The C-ABI library loads and exports every symbol include/realflare_amd.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os

import pytest


def test_library_exports_header_symbols():
    from realflare_amd import native

    if not native.available():
        pytest.fail('librealflare_amd.so is not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(native.LIB_PATH)
    names = native.exported_symbols()
    assert 'rf_trace' in names and 'rf_last_error' in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.rf_abi_version.restype = ctypes.c_int
    assert lib.rf_abi_version() == 2


def test_invalid_arguments_report_error_without_gpu():
    from realflare_amd import native

    lib = native.lib()
    prm = native.TraceParams()
    prm.lens_count = 0
    fn = lib.rf_trace
    fn.restype = ctypes.c_int
    rc = fn(None, None, None, ctypes.byref(prm), None, None)
    assert rc == native.RF_ERR_INVALID
    assert 'lens_count' in native.last_error()
