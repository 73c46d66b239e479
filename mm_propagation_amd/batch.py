"""Batched fibers over the C ABI (mmp_batch_create and the mmp_fiber_* calls).

No reference counterpart: the reference runs one fiber per solver object. A batch holds F
independent fibers that share n_modes, nt, method_order, Raman on/off and the pair-slot
structure of their coupling lists; everything else (dispersion, coupling values, length,
pulse, Raman response) is per fiber. Each stage pass of a step runs all F fibers in one
kernel launch.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _capi
from .problem import CParams, SimulationParameters
from .solver import Context, default_context

_DP = ctypes.POINTER(ctypes.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class MultimodePropagationBatch:
    def __init__(self, params: Sequence[SimulationParameters], raman: bool = True,
                 pair_decode: int = _capi.MMP_PAIRS_PHYSICAL, flags: int = 0,
                 precision: int = _capi.MMP_PRECISION_FP64, ctx: Optional[Context] = None):
        if not params:
            raise ValueError("a batch needs at least one fiber")
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.raman = raman
        self._cps = [CParams(sp, raman) for sp in params]
        arr = (_capi.MmpParams * len(self._cps))(*[cp.c for cp in self._cps])
        self._arr = arr
        opt = _capi.MmpOptions(precision=precision, pair_decode=pair_decode, flags=flags, reserved=0)
        h = ctypes.c_void_p()
        _capi.check(self.lib.mmp_batch_create(self.ctx.h, len(self._cps), ctypes.cast(arr, ctypes.c_void_p),
                                              ctypes.byref(opt), ctypes.byref(h)), "mmp_batch_create")
        self.h = h
        self.F = len(self._cps)
        self.M = int(params[0].n_modes)
        self.nt = int(params[0].nt)

    def computeLawsonRK(self, nz: int) -> None:
        """Every fiber restarts at z = 0 and takes nz steps of its own h."""
        _capi.check(self.lib.mmp_compute_lawson_rk(self.h, int(nz)), "mmp_compute_lawson_rk")

    def advance(self, n: int) -> None:
        _capi.check(self.lib.mmp_advance(self.h, int(n)), "mmp_advance")

    def synchronize(self) -> None:
        self.ctx.synchronize()

    def result(self, f: int) -> np.ndarray:
        out = np.empty((self.M, self.nt), dtype=np.complex128)
        _capi.check(self.lib.mmp_fiber_result(self.h, int(f), _dp(out.view(np.float64))), "mmp_fiber_result")
        return out

    def state(self, f: int):
        psi = np.empty((self.M, self.nt), dtype=np.complex128)
        z = ctypes.c_double()
        _capi.check(self.lib.mmp_fiber_state(self.h, int(f), _dp(psi.view(np.float64)), ctypes.byref(z)),
                    "mmp_fiber_state")
        return psi, z.value

    def set_state(self, f: int, psi: np.ndarray) -> None:
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        if psi.shape != (self.M, self.nt):
            raise ValueError("psi must have shape (M, NT)")
        _capi.check(self.lib.mmp_fiber_set_state(self.h, int(f), _dp(psi.view(np.float64))), "mmp_fiber_set_state")

    def energy(self, f: int) -> np.ndarray:
        out = np.empty(self.M, dtype=np.float64)
        _capi.check(self.lib.mmp_fiber_energy(self.h, int(f), _dp(out)), "mmp_fiber_energy")
        return out

    def photon_number(self, f: int, wavelength: float) -> np.ndarray:
        """Fiber f's per-mode photon number for a carrier wavelength in metres."""
        out = np.empty(self.M, dtype=np.float64)
        _capi.check(self.lib.mmp_fiber_photon_number(self.h, int(f), float(wavelength), _dp(out)),
                    "mmp_fiber_photon_number")
        return out

    def info(self, f: int = 0) -> Dict[str, float]:
        i = _capi.MmpInfo()
        _capi.check(self.lib.mmp_fiber_info(self.h, int(f), ctypes.byref(i)), "mmp_fiber_info")
        return {name: getattr(i, name) for name, _ in _capi.MmpInfo._fields_}

    def load(self, f: int, params: SimulationParameters) -> None:
        """Replace fiber f by a new problem of the same structure; it restarts at z = 0."""
        cp = CParams(params, self.raman)
        _capi.check(self.lib.mmp_fiber_load(self.h, int(f), ctypes.byref(cp.c)), "mmp_fiber_load")
        self._cps[f] = cp

    def set_live(self, f: int, live: bool) -> None:
        """Park (False) or resume (True) fiber f: a parked fiber is not advanced by the steps and
        its workgroups exit at once."""
        _capi.check(self.lib.mmp_fiber_set_live(self.h, int(f), 1 if live else 0), "mmp_fiber_set_live")

    def kernel_times(self) -> List[dict]:
        n = ctypes.c_uint32()
        _capi.check(self.lib.mmp_kernel_times(self.h, None, 0, ctypes.byref(n)), "mmp_kernel_times")
        arr = (_capi.MmpKernelTime * max(n.value, 1))()
        _capi.check(self.lib.mmp_kernel_times(self.h, arr, n.value, ctypes.byref(n)), "mmp_kernel_times")
        return [{"name": arr[i].name.decode(), "total_ms": arr[i].total_ms, "launches": int(arr[i].launches),
                 "bytes_per_launch": arr[i].bytes_per_launch} for i in range(n.value)]

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mmp_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SolverGroup:
    """Solvers and batches of different shapes on one context, stepped together by one HIP graph
    (mmp_group_*): advance(n) equals advance(n) on every member, bit for bit."""

    def __init__(self, members: Sequence, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.lib
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _capi.check(self.lib.mmp_group_create(ctx.h, len(self.members), arr, ctypes.byref(h)), "mmp_group_create")
        self.h = h

    def advance(self, n: int) -> None:
        _capi.check(self.lib.mmp_group_advance(self.h, int(n)), "mmp_group_advance")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mmp_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
