"""ctypes binding of ``_lib/librealflare_amd.so`` (the C ABI in
``include/realflare_amd.h``).

torch is imported before the library is opened, so the library binds to the
HIP runtime torch already loaded and device pointers from torch tensors are
valid in its calls.  Every call enqueues on torch's current stream.  There is
no fallback: without the built library every render call raises
``NativeLibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
from functools import lru_cache

import torch  # noqa: F401  (must be loaded before the HIP library)

from .api.data import RealflareError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_lib', 'librealflare_amd.so')

RF_OK, RF_ERR_INVALID, RF_ERR_HIP, RF_ERR_WORKSPACE = 0, 1, 2, 3


class NativeLibraryMissing(RuntimeError):
    pass


class NativeError(RuntimeError):
    pass


class TraceParams(ctypes.Structure):
    _fields_ = [
        ('lens_count', ctypes.c_int32),
        ('path_count', ctypes.c_int32),
        ('wavelength_count', ctypes.c_int32),
        ('aperture_index', ctypes.c_int32),
        ('coating_min_ior', ctypes.c_float),
        ('grid_count', ctypes.c_int32),
        ('grid_length', ctypes.c_float),
        ('direction', ctypes.c_float * 4),
    ]


@lru_cache(1)
def lib() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f'{LIB_PATH} is not built; run realflare_amd/build_native.py (or '
            '__graft_entry__.build()) first'
        )
    handle = ctypes.CDLL(LIB_PATH)
    handle.rf_last_error.restype = ctypes.c_char_p
    return handle


def available() -> bool:
    return os.path.exists(LIB_PATH)


def last_error() -> str:
    return lib().rf_last_error().decode(errors='replace')


def call(name: str, *args) -> None:
    fn = getattr(lib(), name)
    fn.restype = ctypes.c_int
    rc = fn(*args)
    if rc == RF_OK:
        return
    msg = f'{name}: {last_error()}'
    if rc == RF_ERR_INVALID:
        raise RealflareError(msg)
    raise NativeError(msg)


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_cuda:
        raise ValueError('device tensor expected')
    if not t.is_contiguous():
        raise ValueError('contiguous tensor expected')
    return ctypes.c_void_p(t.data_ptr())


def ptr_at(t, offset_bytes: int) -> ctypes.c_void_p:
    """Device address `offset_bytes` into tensor `t`."""
    if not 0 <= offset_bytes <= t.numel() * t.element_size():
        raise ValueError('offset outside the tensor')
    return ctypes.c_void_p(ptr(t).value + int(offset_bytes))


def exported_symbols() -> list[str]:
    """Names declared in include/realflare_amd.h (parsed from the header)."""
    import re

    header = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include',
                          'realflare_amd.h')
    text = open(header).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(rf_[a-z0-9_]+)\s*\(', text)))
