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
    assert lib.rf_abi_version() == 3


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


def test_raster_ghost_texture_limits_without_gpu():
    """The rasterizer reads the ghost texture through a zero-framed copy with
    24-bit texel arithmetic: each side must stay below 2^24 - 4 texels (and the
    texture at most 2^28 texels); the size query answers 0 and the entry point
    RF_ERR_INVALID beyond that."""
    from realflare_amd import native
    from realflare_amd.api.tasks.rasterizing import raster_params

    lib = native.lib()
    size = lib.rf_rasterize_workspace_size
    size.restype = ctypes.c_size_t
    prm = raster_params(2, 2, 4, 1.0, (64, 64), 64, 1, 1, (36.0, 24.0), 0.0, 1.0, 8.0, (16, 16), 4)
    ok = size(ctypes.byref(prm), ctypes.c_uint64(1024))
    assert ok > 0
    # one more texel per side grows the workspace (the framed copy is in it)
    prm.ghost_width, prm.ghost_height = 4096, 4096
    assert size(ctypes.byref(prm), ctypes.c_uint64(1024)) >= ok + 4096 * 4096 * 4
    prm.ghost_width, prm.ghost_height = (1 << 24) - 5, 1
    assert size(ctypes.byref(prm), ctypes.c_uint64(1024)) > 0
    prm.ghost_width = (1 << 24) - 4
    assert size(ctypes.byref(prm), ctypes.c_uint64(1024)) == 0
    fn = lib.rf_raster_bin
    fn.restype = ctypes.c_int
    rc = fn(None, None, None, ctypes.byref(prm), ctypes.c_uint64(0), None, ctypes.c_size_t(0), None)
    assert rc == native.RF_ERR_INVALID
    assert 'ghost texture side' in native.last_error()


def test_starburst_spectrum_limits_without_gpu():
    """The starburst reads the power spectrum through a raw buffer with 32-bit
    byte offsets and clamps texel coordinates as floats: at most 2^28 values
    and sides below 2^24, RF_ERR_INVALID beyond that (checked before any
    launch; the pointers are never read)."""
    from realflare_amd import native
    from realflare_amd.api.data import Flare
    from realflare_amd.api.tasks.starburst import starburst_params

    lib = native.lib()
    fn = lib.rf_starburst
    fn.restype = ctypes.c_int
    prm = starburst_params(Flare.Starburst(), (16, 16), 8, (0.0, 0.0), (1.0, 1.0))
    buf = ctypes.create_string_buffer(64)
    for w, h in [((1 << 14) + 1, 1 << 14), (1 << 24, 1), (1, 1 << 24)]:
        rc = fn(buf, w, h, buf, 4, ctypes.byref(prm), buf, None)
        assert rc == native.RF_ERR_INVALID, (w, h)
        assert 'power spectrum' in native.last_error()
    prm.samples = 0
    rc = fn(buf, 4, 4, buf, 4, ctypes.byref(prm), buf, None)
    assert rc == native.RF_ERR_INVALID and 'samples' in native.last_error()


def test_flare_static_rejects_frames_beyond_its_scratch_without_gpu():
    """rf_flare_static's workspace is a fixed device array: a frame whose
    rf_flare workspace exceeds it even with no binning pairs is refused
    before anything touches the device, with the sizes in the message."""
    from realflare_amd import native
    from realflare_amd.api.tasks.rasterizing import raster_params

    lib = native.lib()
    size = lib.rf_flare_workspace_size
    size.restype = ctypes.c_size_t
    limit = 256 << 20
    # 66 ghosts x 8 wavelengths x 128^2 rays at 2048^2: far above 256 MiB
    big = raster_params(66, 8, 128, 50.0, (2048, 2048), 64, 1, 1, (36.0, 24.0), 0.01, 1.0, 8.0,
                        (256, 256), 340)
    assert size(ctypes.byref(big), ctypes.c_uint64(0)) > limit
    fn = lib.rf_flare_static
    fn.restype = ctypes.c_int
    rc = fn(None, None, None, ctypes.byref(big), None, None, None)
    assert rc == native.RF_ERR_WORKSPACE
    assert 'static scratch' in native.last_error() and str(limit) in native.last_error()
    # a small frame fits (its NULL buffers are then the error)
    small = raster_params(15, 3, 64, 50.0, (512, 512), 64, 1, 1, (36.0, 24.0), 0.01, 1.0, 8.0,
                          (128, 128), 340)
    assert size(ctypes.byref(small), ctypes.c_uint64(1 << 20)) < limit
    rc = fn(None, None, None, ctypes.byref(small), None, None, None)
    assert rc == native.RF_ERR_INVALID and 'NULL' in native.last_error()
