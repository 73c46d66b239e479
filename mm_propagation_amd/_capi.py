"""ctypes view of the C ABI declared in include/mmprop.h (libmmprop_hip.so).

The structures mirror the C declarations field for field. Loading the library
is explicit (`load_library`); nothing here falls back to another implementation:
if the HIP library is missing, `load_library` raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_int, c_uint32, c_uint64, c_size_t, c_void_p

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libmmprop_hip.so"
LIB_PATH = os.path.join(PKG_DIR, "lib", LIB_NAME)

MMP_OK = 0
MMP_ERR_INVALID_ARGUMENT = 1
MMP_ERR_UNSUPPORTED = 2
MMP_ERR_NONFINITE = 6
MMP_ERR_STATE = 7
STATUS_NAMES = {
    0: "MMP_OK",
    1: "MMP_ERR_INVALID_ARGUMENT",
    2: "MMP_ERR_UNSUPPORTED",
    3: "MMP_ERR_OUT_OF_MEMORY",
    4: "MMP_ERR_DEVICE",
    5: "MMP_ERR_NO_DEVICE",
    6: "MMP_ERR_NONFINITE",
    7: "MMP_ERR_STATE",
}

MMP_PRECISION_FP64 = 0
MMP_PRECISION_MIXED = 1
MMP_SNAPSHOTS_REFERENCE_RESTART = 0
MMP_SNAPSHOTS_RESUMABLE = 1
MMP_PAIRS_REFERENCE_CPU = 0
MMP_PAIRS_PHYSICAL = 1
MMP_FLAG_PROFILE = 1 << 0
MMP_FLAG_NO_GRAPH = 1 << 1
MMP_FLAG_SEPARATE_COLUMNS = 1 << 2
MMP_FLAG_AUTOTUNE = 1 << 3
MMP_FLAG_NO_MATRIX_CORES = 1 << 4
MMP_FLAG_MATRIX_CORES = 1 << 5


class MmpParams(ctypes.Structure):
    _fields_ = [
        ("n_modes", c_uint32),
        ("fiber_length", c_double),
        ("n_beta", c_uint32),
        ("dispersion_coefficients", POINTER(c_double)),
        ("coupling_offsets", POINTER(c_uint32)),
        ("coupling_klm", POINTER(c_uint32)),
        ("coupling_values", POINTER(c_double)),
        ("nt", c_uint32),
        ("time_window", c_double),
        ("pulse_width", c_double),
        ("initial_fields", POINTER(c_double)),
        ("n_steps", c_uint32),
        ("method_order", c_uint32),
        ("nonlinearity_const", c_double),
        ("raman_proportion", c_double),
        ("raman_response", POINTER(c_double)),
        ("raman", c_uint32),
    ]


class MmpOptions(ctypes.Structure):
    _fields_ = [
        ("precision", c_uint32),
        ("pair_decode", c_uint32),
        ("flags", c_uint32),
        ("reserved", c_uint32),
    ]


class MmpInfo(ctypes.Structure):
    _fields_ = [
        ("n_modes", c_uint32), ("nt", c_uint32), ("stages", c_uint32), ("n_pairs", c_uint32),
        ("nnz", c_uint32), ("n_conv_lines", c_uint32), ("n1", c_uint32), ("n2", c_uint32),
        ("steps_taken", c_uint32), ("default_steps", c_uint32),
        ("t_final", c_double), ("delta_t", c_double), ("z_final", c_double), ("delta_z", c_double),
        ("z", c_double), ("conversion_factor", c_double), ("sign_gamma", c_double), ("ld", c_double),
        ("max_beta2", c_double), ("max_q", c_double),
        ("device_bytes", c_size_t),
        ("n_fibers", c_uint32), ("live", c_uint32),
        ("gather_waves", c_uint32), ("chunk_lines", c_uint32), ("tuned", c_uint32),
        ("pair_columns", c_uint32), ("pair_groups", c_uint32),
        ("gather_hi_occupancy", c_uint32), ("gather_matrix_cores", c_uint32),
        ("gather_chunks", c_uint32), ("halted", c_uint32),
    ]


MMP_REQ_QUEUED = 0
MMP_REQ_RUNNING = 1
MMP_REQ_DONE = 2
MMP_REQ_NONFINITE = 3


class MmpServeOptions(ctypes.Structure):
    _fields_ = [
        ("device", c_int), ("slots", c_uint32), ("max_chunk", c_uint32), ("precision", c_uint32),
        ("pair_decode", c_uint32), ("flags", c_uint32), ("hbm_budget", c_double),
    ]


class MmpServeRecord(ctypes.Structure):
    _fields_ = [
        ("state", c_uint32), ("batched", c_uint32), ("group", c_uint32), ("slot", c_uint32),
        ("t_submit", c_double), ("t_admit", c_double), ("t_finish", c_double),
        ("units", c_uint64), ("device_bytes", c_size_t),
    ]


class MmpServeStats(ctypes.Structure):
    _fields_ = [
        ("submitted", c_uint64), ("queued", c_uint64), ("running", c_uint64), ("done", c_uint64),
        ("nonfinite", c_uint64), ("groups", c_uint32), ("solos", c_uint32), ("groups_created", c_uint32),
        ("solos_created", c_uint32), ("bytes_in_use", c_size_t), ("peak_bytes", c_size_t),
    ]


class MmpKernelTime(ctypes.Structure):
    _fields_ = [
        ("name", c_char * 48),
        ("total_ms", c_double),
        ("launches", c_uint64),
        ("bytes_per_launch", c_double),
    ]


# (name, restype, argtypes) of every exported entry point of include/mmprop.h
EXPORTS = [
    ("mmp_status_string", c_char_p, [c_int]),
    ("mmp_last_error", c_char_p, []),
    ("mmp_abi_version", c_int, []),
    ("mmp_source_hash", c_char_p, []),
    ("mmp_ctx_create", c_int, [c_int, POINTER(c_void_p)]),
    ("mmp_ctx_destroy", None, [c_void_p]),
    ("mmp_solver_create", c_int, [c_void_p, POINTER(MmpParams), POINTER(MmpOptions), POINTER(c_void_p)]),
    ("mmp_solver_destroy", None, [c_void_p]),
    ("mmp_compute_lawson_rk", c_int, [c_void_p, c_uint32]),
    ("mmp_compute_lawson_rk_default", c_int, [c_void_p]),
    ("mmp_advance", c_int, [c_void_p, c_uint32]),
    ("mmp_get_result", c_int, [c_void_p, POINTER(c_double)]),
    ("mmp_batch_create", c_int, [c_void_p, c_uint32, c_void_p, c_void_p, POINTER(c_void_p)]),
    ("mmp_ctx_query", c_int, [c_void_p]),
    ("mmp_group_create", c_int, [c_void_p, c_uint32, POINTER(c_void_p), POINTER(c_void_p)]),
    ("mmp_group_advance", c_int, [c_void_p, c_uint32]),
    ("mmp_group_destroy", None, [c_void_p]),
    ("mmp_solve_static", c_int, [c_void_p, POINTER(MmpParams), POINTER(MmpOptions), c_uint32, POINTER(c_double)]),
    ("mmp_fiber_result", c_int, [c_void_p, c_uint32, POINTER(c_double)]),
    ("mmp_fiber_state", c_int, [c_void_p, c_uint32, POINTER(c_double), POINTER(c_double)]),
    ("mmp_fiber_set_state", c_int, [c_void_p, c_uint32, POINTER(c_double)]),
    ("mmp_fiber_energy", c_int, [c_void_p, c_uint32, POINTER(c_double)]),
    ("mmp_fiber_info", c_int, [c_void_p, c_uint32, c_void_p]),
    ("mmp_fiber_load", c_int, [c_void_p, c_uint32, c_void_p]),
    ("mmp_fiber_set_live", c_int, [c_void_p, c_uint32, c_uint32]),
    ("mmp_propagate_snapshots", c_int, [c_void_p, c_uint32, c_uint32, c_uint32, POINTER(c_double),
                                        POINTER(c_double)]),
    ("mmp_get_state", c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
    ("mmp_set_state", c_int, [c_void_p, POINTER(c_double)]),
    ("mmp_energy", c_int, [c_void_p, POINTER(c_double)]),
    ("mmp_photon_number", c_int, [c_void_p, c_double, POINTER(c_double)]),
    ("mmp_fiber_photon_number", c_int, [c_void_p, c_uint32, c_double, POINTER(c_double)]),
    ("mmp_solver_info", c_int, [c_void_p, POINTER(MmpInfo)]),
    ("mmp_kernel_times", c_int, [c_void_p, POINTER(MmpKernelTime), c_uint32, POINTER(c_uint32)]),
    ("mmp_ctx_synchronize", c_int, [c_void_p]),
    ("mmp_serve_create", c_int, [POINTER(MmpServeOptions), POINTER(c_void_p)]),
    ("mmp_serve_destroy", None, [c_void_p]),
    ("mmp_serve_submit", c_int, [c_void_p, POINTER(MmpParams), c_uint32, POINTER(c_uint64)]),
    ("mmp_serve_poll", c_int, [c_void_p, POINTER(c_uint32)]),
    ("mmp_serve_drain", c_int, [c_void_p]),
    ("mmp_serve_record_get", c_int, [c_void_p, c_uint64, POINTER(MmpServeRecord)]),
    ("mmp_serve_result", c_int, [c_void_p, c_uint64, POINTER(c_double)]),
    ("mmp_serve_stats_get", c_int, [c_void_p, POINTER(MmpServeStats)]),
    ("mmp_plan", c_int, [POINTER(MmpParams), POINTER(MmpOptions), POINTER(MmpInfo)]),
    ("mmp_tableau", c_int, [c_uint32, POINTER(c_uint32), POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    ("mmp_plan_arrays", c_int, [POINTER(MmpParams), POINTER(MmpOptions), POINTER(c_double), POINTER(c_double),
                                POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), POINTER(c_double)]),
]

_LIB = None


class MmpError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, f"status {status}")
        super().__init__(f"{where}: {name}" + (f" ({detail})" if detail else ""))


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libmmprop_hip.so (built in-tree by mm_propagation_amd.build). Raises if it is absent,
    or if it was built from other sources than the tree's (its mmp_source_hash differs from the
    digest of the sources here)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    for name, restype, argtypes in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    if path is None:
        from . import build
        built, here = lib.mmp_source_hash().decode(), build.source_hash()
        if built != here:
            raise RuntimeError(f"{p} was built from other sources (digest {built}, the tree's sources {here}): "
                               "rebuild it (python -c 'import __graft_entry__ as g; g.build()')")
        _LIB = lib
    return lib


def check(status: int, where: str) -> None:
    if status != MMP_OK:
        detail = ""
        if _LIB is not None:
            raw = _LIB.mmp_last_error()
            detail = raw.decode() if raw else ""
        raise MmpError(status, where, detail)
