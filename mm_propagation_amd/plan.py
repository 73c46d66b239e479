"""Host-only planning through the C ABI (no device needed): validation, solver scalars,
device-memory footprint and the prepared arrays (mmp_plan / mmp_plan_arrays)."""
from __future__ import annotations

import ctypes
from typing import Dict

import numpy as np

from . import _capi
from .problem import CParams, SimulationParameters

_DP = ctypes.POINTER(ctypes.c_double)
_UP = ctypes.POINTER(ctypes.c_uint32)


def _opts(pair_decode: int, precision: int = _capi.MMP_PRECISION_FP64, flags: int = 0):
    return _capi.MmpOptions(precision=precision, pair_decode=pair_decode, flags=flags, reserved=0)


def plan(sp: SimulationParameters, raman: bool = True, pair_decode: int = 0,
         precision: int = _capi.MMP_PRECISION_FP64, flags: int = 0) -> Dict[str, float]:
    """Scalars and device bytes a solver for `sp` would have; raises MmpError on invalid input."""
    lib = _capi.load_library()
    cp = CParams(sp, raman)
    info = _capi.MmpInfo()
    o = _opts(pair_decode, precision, flags)
    _capi.check(lib.mmp_plan(ctypes.byref(cp.c), ctypes.byref(o), ctypes.byref(info)), "mmp_plan")
    return {name: getattr(info, name) for name, _ in _capi.MmpInfo._fields_}


def plan_arrays(sp: SimulationParameters, raman: bool = True, pair_decode: int = 0) -> Dict[str, np.ndarray]:
    lib = _capi.load_library()
    info = plan(sp, raman, pair_decode)
    cp = CParams(sp, raman)
    M, nt, P, s = int(sp.n_modes), int(sp.nt), int(info["n_pairs"]), int(info["stages"])
    psi0 = np.zeros((M, nt), np.complex128)
    imL = np.zeros((M, nt))
    pa = np.zeros(max(P, 1), np.uint32)
    pb = np.zeros(max(P, 1), np.uint32)
    e1 = np.zeros((M, nt), np.complex128)
    e = np.zeros((s, M, nt), np.complex128)
    o = _opts(pair_decode)
    _capi.check(lib.mmp_plan_arrays(ctypes.byref(cp.c), ctypes.byref(o), psi0.view(np.float64).ctypes.data_as(_DP),
                                    imL.ctypes.data_as(_DP), pa.ctypes.data_as(_UP), pb.ctypes.data_as(_UP),
                                    e1.view(np.float64).ctypes.data_as(_DP), e.view(np.float64).ctypes.data_as(_DP)),
                "mmp_plan_arrays")
    return {"psi0": psi0, "imL": imL, "pair_a": pa[:P], "pair_b": pb[:P], "e1": e1, "e": e, "info": info}
