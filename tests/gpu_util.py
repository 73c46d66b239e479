"""Helpers shared by the GPU tests (device checks, solver construction)."""
import os

import numpy as np
import pytest

from mm_propagation_amd import _capi


def device_available() -> bool:
    try:
        lib = _capi.load_library()
    except Exception:
        return False
    import ctypes
    h = ctypes.c_void_p()
    st = lib.mmp_ctx_create(0, ctypes.byref(h))
    if st == 0:
        lib.mmp_ctx_destroy(h)
        return True
    return False


def require_device():
    if not device_available():
        pytest.fail("GPU test needs libmmprop_hip.so and a gfx950 device (the HIP path has no fallback)")


def note(msg: str) -> None:
    """Name the case before a GPU step, flushed, so a failing run says what it was doing."""
    print(msg, flush=True)
