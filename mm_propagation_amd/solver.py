"""Python mirror of the reference's solver classes over the C ABI.

`MultimodePropagationGPU_RamanON` / `MultimodePropagationGPU_RamanOFF` keep the
names, constructor argument and methods of include/MultimodePropagation_GPU.cuh:182-213
(computeLawsonRK(nz), computeLawsonRK(), getResult()). All propagation runs in
libmmprop_hip.so on a gfx950 device; without the library or a device the
constructor raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional

import numpy as np

from . import _capi
from .problem import CParams, SimulationParameters


class Context:
    """One HIP stream on one device (mmp_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = _capi.load_library()
        h = ctypes.c_void_p()
        _capi.check(self.lib.mmp_ctx_create(device, ctypes.byref(h)), "mmp_ctx_create")
        self.h = h
        self.device = device

    def synchronize(self) -> None:
        _capi.check(self.lib.mmp_ctx_synchronize(self.h), "mmp_ctx_synchronize")

    def idle(self) -> bool:
        """True when all work queued on this context's stream has completed (non-blocking)."""
        st = self.lib.mmp_ctx_query(self.h)
        if st == _capi.MMP_ERR_STATE:
            return False
        _capi.check(st, "mmp_ctx_query")
        return True

    def close(self) -> None:
        if self.h:
            self.lib.mmp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class MultimodePropagationGPU:
    """Base of the two drop-in classes; `raman` picks the class behaviour."""

    raman = True

    def __init__(self, params: SimulationParameters, pair_decode: int = _capi.MMP_PAIRS_PHYSICAL,
                 flags: int = 0, ctx: Optional[Context] = None, precision: int = _capi.MMP_PRECISION_FP64):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.params = params
        self._cp = CParams(params, self.raman)
        opt = _capi.MmpOptions(precision=precision, pair_decode=pair_decode, flags=flags, reserved=0)
        h = ctypes.c_void_p()
        _capi.check(self.lib.mmp_solver_create(self.ctx.h, ctypes.byref(self._cp.c), ctypes.byref(opt),
                                               ctypes.byref(h)), "mmp_solver_create")
        self.h = h
        self.M = int(params.n_modes)
        self.nt = int(params.nt)

    # --- reference API -------------------------------------------------
    def computeLawsonRK(self, nz: Optional[int] = None) -> None:
        """computeLawsonRK(nz) (restart at z = 0, keep psi) or computeLawsonRK() with
        nz = floor(z_final / delta_z)."""
        if nz is None:
            _capi.check(self.lib.mmp_compute_lawson_rk_default(self.h), "mmp_compute_lawson_rk_default")
        else:
            _capi.check(self.lib.mmp_compute_lawson_rk(self.h, int(nz)), "mmp_compute_lawson_rk")

    def getResult(self) -> np.ndarray:
        out = np.empty((self.M, self.nt), dtype=np.complex128)
        _capi.check(self.lib.mmp_get_result(self.h, out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "mmp_get_result")
        return out

    # --- extensions ----------------------------------------------------
    def propagate_snapshots(self, n_steps: int, n_save: int, resumable: bool = False):
        """The MEX gateway's n_save loop (matlab/mm_propagation.cpp:61-101): returns the physical
        fields [n_save][M][NT] and the distances Z (m) after each chunk. resumable=False restarts
        z and the Lawson tables per chunk as the gateway does; True runs one solve paused at each
        snapshot."""
        fields = np.empty((n_save, self.M, self.nt), dtype=np.complex128)
        z = np.empty(n_save, dtype=np.float64)
        mode = _capi.MMP_SNAPSHOTS_RESUMABLE if resumable else _capi.MMP_SNAPSHOTS_REFERENCE_RESTART
        _capi.check(self.lib.mmp_propagate_snapshots(self.h, int(n_steps), int(n_save), mode,
                                                     fields.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "mmp_propagate_snapshots")
        return fields, z

    def advance(self, n: int) -> None:
        """Resumable stepping from the current z (no restart)."""
        _capi.check(self.lib.mmp_advance(self.h, int(n)), "mmp_advance")

    def get_state(self):
        psi = np.empty((self.M, self.nt), dtype=np.complex128)
        z = ctypes.c_double()
        _capi.check(self.lib.mmp_get_state(self.h, psi.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           ctypes.byref(z)), "mmp_get_state")
        return psi, z.value

    def set_state(self, psi: np.ndarray) -> None:
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        if psi.shape != (self.M, self.nt):
            raise ValueError("psi must have shape (M, NT)")
        _capi.check(self.lib.mmp_set_state(self.h, psi.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "mmp_set_state")

    def energy(self) -> np.ndarray:
        out = np.empty(self.M, dtype=np.float64)
        _capi.check(self.lib.mmp_energy(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "mmp_energy")
        return out

    def photon_number(self, wavelength: float) -> np.ndarray:
        """Per-mode photon number of the physical field for a carrier wavelength in metres."""
        out = np.empty(self.M, dtype=np.float64)
        _capi.check(self.lib.mmp_photon_number(self.h, float(wavelength),
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "mmp_photon_number")
        return out

    def info(self) -> Dict[str, float]:
        i = _capi.MmpInfo()
        _capi.check(self.lib.mmp_solver_info(self.h, ctypes.byref(i)), "mmp_solver_info")
        return {name: getattr(i, name) for name, _ in _capi.MmpInfo._fields_}

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


class MultimodePropagationGPU_RamanON(MultimodePropagationGPU):
    raman = True


class MultimodePropagationGPU_RamanOFF(MultimodePropagationGPU):
    raman = False


def make_solver(params: SimulationParameters, **kw) -> MultimodePropagationGPU:
    """Class choice of matlab/mm_propagation.cpp:59: Raman off when raman_proportion == 0."""
    cls = MultimodePropagationGPU_RamanOFF if params.raman_proportion == 0.0 else MultimodePropagationGPU_RamanON
    return cls(params, **kw)


def solve_static(params: SimulationParameters, nz: int, raman: bool = True,
                 pair_decode: int = _capi.MMP_PAIRS_PHYSICAL, precision: int = _capi.MMP_PRECISION_FP64,
                 flags: int = 0, ctx: Optional[Context] = None) -> np.ndarray:
    """mmp_solve_static: computeLawsonRK(nz) from the initial field and the physical field at the
    end, with every device array in the library's static per-device scratch (no device
    allocation). Same result as the solver classes, bit for bit."""
    ctx = ctx or default_context()
    cp = CParams(params, raman)
    opt = _capi.MmpOptions(precision=precision, pair_decode=pair_decode, flags=flags, reserved=0)
    out = np.empty((int(params.n_modes), int(params.nt)), dtype=np.complex128)
    _capi.check(ctx.lib.mmp_solve_static(ctx.h, ctypes.byref(cp.c), ctypes.byref(opt), int(nz),
                                         out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                "mmp_solve_static")
    return out
