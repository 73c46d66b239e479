"""ctypes binding of the checker libraries (oracle/_build/libmmprop_oracle.so and,
where it was built, oracle/_ref/libref_setup.so). Test infrastructure only."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_uint, c_void_p

import numpy as np

from mm_propagation_amd import _capi
from mm_propagation_amd.problem import CParams, SimulationParameters

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libmmprop_oracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libref_setup.so")


class OrcOptions(ctypes.Structure):
    _fields_ = [("pair_decode", c_int), ("fft_variant", c_int), ("parallel_pairs", c_int), ("n_threads", c_int),
                ("fp32_stages", c_int)]


# ORC_FP32_* of oracle/oracle.h: stages whose arithmetic an error-budget study rounds to FP32
ORC_FP32_FIELDS, ORC_FP32_PAIRS, ORC_FP32_RAMAN, ORC_FP32_GATHER, ORC_FP32_SIGMA = 1, 2, 4, 8, 16


class OrcSetupInfo(ctypes.Structure):
    _fields_ = [("M", c_uint), ("nt", c_uint), ("stages", c_uint), ("n_pairs", c_uint), ("n_beta", c_uint),
                ("t_final", c_double), ("delta_t", c_double), ("z_final", c_double), ("delta_z", c_double),
                ("z", c_double), ("conversion_factor", c_double), ("sign_gamma", c_double), ("T0", c_double),
                ("ld", c_double), ("max_beta2", c_double), ("max_q", c_double)]


_DP = POINTER(c_double)
_UP = POINTER(c_uint)
_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        sigs = {
            "orc_create": (c_void_p, [POINTER(_capi.MmpParams), POINTER(OrcOptions)]),
            "orc_destroy": (None, [c_void_p]),
            "orc_last_error": (c_char_p, []),
            "orc_compute_lawson_rk": (None, [c_void_p, c_uint]),
            "orc_advance": (None, [c_void_p, c_uint]),
            "orc_default_steps": (c_uint, [c_void_p]),
            "orc_get_result": (None, [c_void_p, _DP]),
            "orc_get_state": (None, [c_void_p, _DP, _DP]),
            "orc_set_state": (None, [c_void_p, _DP]),
            "orc_get_setup": (None, [c_void_p, POINTER(OrcSetupInfo)]),
            "orc_get_tableau": (None, [c_void_p, _DP, _DP, _DP]),
            "orc_get_beta_tilde": (None, [c_void_p, _DP]),
            "orc_get_q_tilde": (None, [c_void_p, _DP]),
            "orc_get_unique_pairs": (None, [c_void_p, _UP]),
            "orc_get_pair_table": (None, [c_void_p, _UP, _UP]),
            "orc_get_nnz_slots": (None, [c_void_p, _UP]),
            "orc_init_operators": (None, [c_void_p]),
            "orc_get_linear_operator": (None, [c_void_p, _DP]),
            "orc_get_e1": (None, [c_void_p, _DP]),
            "orc_get_e": (None, [c_void_p, c_uint, _DP]),
            "orc_get_hr_fft": (None, [c_void_p, _DP]),
            "orc_nonlinearity": (None, [c_void_p, c_uint, _DP, _DP]),
            "orc_dft": (None, [c_uint, _DP, _DP, c_int, c_int]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_SO)
        f = L.ref_probe_setup
        f.restype = c_int
        f.argtypes = [POINTER(_capi.MmpParams), _UP, _DP, _DP, _DP, _DP, _DP, _DP, _UP, _DP, _DP, _DP, _DP]
        _ref = L
    return _ref


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def _up(a: np.ndarray):
    return a.ctypes.data_as(_UP)


def dft(x: np.ndarray, sign: int, variant: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_dft(x.size, _dp(x.view(np.float64)), _dp(out.view(np.float64)), sign, variant)
    return out


class Oracle:
    """The C restatement of the reference CPU solver."""

    def __init__(self, sp: SimulationParameters, raman: bool = True, pair_decode: int = 0, fft_variant: int = 0,
                 parallel_pairs: bool = True, n_threads: int = -1, fp32_stages: int = 0):
        self.sp = sp
        self.cp = CParams(sp, raman)
        self.M = int(sp.n_modes)
        self.nt = int(sp.nt)
        self.nnz = self.cp.nnz
        if n_threads < 0:
            # small problems run single-threaded: thousands of tiny OpenMP regions per solve
            # cost more in barriers than they gain
            work = int(sp.n_modes) * int(sp.nt)
            n_threads = 1 if work < 65536 else min(8, os.cpu_count() or 1)
        opt = OrcOptions(pair_decode, fft_variant, 1 if parallel_pairs else 0, n_threads, fp32_stages)
        self.h = lib().orc_create(ctypes.byref(self.cp.c), ctypes.byref(opt))
        if not self.h:
            raise ValueError(lib().orc_last_error().decode())
        self.info = OrcSetupInfo()
        lib().orc_get_setup(self.h, ctypes.byref(self.info))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _cbuf(self, n=None):
        return np.zeros((self.M if n is None else n, self.nt), dtype=np.complex128)

    def compute_lawson_rk(self, nz: int):
        lib().orc_compute_lawson_rk(self.h, nz)

    def advance(self, nz: int):
        lib().orc_advance(self.h, nz)

    def default_steps(self) -> int:
        return int(lib().orc_default_steps(self.h))

    def get_result(self) -> np.ndarray:
        out = self._cbuf()
        lib().orc_get_result(self.h, _dp(out.view(np.float64)))
        return out

    def get_state(self):
        out = self._cbuf()
        z = c_double()
        lib().orc_get_state(self.h, _dp(out.view(np.float64)), ctypes.byref(z))
        return out, z.value

    def set_state(self, psi: np.ndarray):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        lib().orc_set_state(self.h, _dp(psi.view(np.float64)))

    def tableau(self):
        s = self.info.stages
        a = np.zeros(s * s)
        b = np.zeros(s)
        c = np.zeros(s)
        lib().orc_get_tableau(self.h, _dp(a), _dp(b), _dp(c))
        return a.reshape(s, s), b, c

    def beta_tilde(self):
        out = np.zeros(self.M * self.info.n_beta)
        lib().orc_get_beta_tilde(self.h, _dp(out))
        return out.reshape(self.M, -1)

    def q_tilde(self):
        out = np.zeros(max(self.nnz, 1))
        lib().orc_get_q_tilde(self.h, _dp(out))
        return out[: self.nnz]

    def unique_pairs(self):
        out = np.zeros(max(self.info.n_pairs, 1), dtype=np.uint32)
        lib().orc_get_unique_pairs(self.h, _up(out))
        return out[: self.info.n_pairs]

    def pair_table(self):
        P = max(self.info.n_pairs, 1)
        a = np.zeros(P, dtype=np.uint32)
        b = np.zeros(P, dtype=np.uint32)
        lib().orc_get_pair_table(self.h, _up(a), _up(b))
        return a[: self.info.n_pairs], b[: self.info.n_pairs]

    def nnz_slots(self):
        out = np.zeros(max(self.nnz, 1), dtype=np.uint32)
        lib().orc_get_nnz_slots(self.h, _up(out))
        return out[: self.nnz]

    def init_operators(self):
        lib().orc_init_operators(self.h)

    def linear_operator(self):
        out = self._cbuf()
        lib().orc_get_linear_operator(self.h, _dp(out.view(np.float64)))
        return out

    def e1(self):
        out = self._cbuf()
        lib().orc_get_e1(self.h, _dp(out.view(np.float64)))
        return out

    def e(self, stage: int):
        out = self._cbuf()
        lib().orc_get_e(self.h, stage, _dp(out.view(np.float64)))
        return out

    def hr_fft(self):
        out = np.zeros(self.nt, dtype=np.complex128)
        lib().orc_get_hr_fft(self.h, _dp(out.view(np.float64)))
        return out

    def nonlinearity(self, stage: int, psi_i: np.ndarray) -> np.ndarray:
        psi_i = np.ascontiguousarray(psi_i, dtype=np.complex128)
        out = self._cbuf()
        lib().orc_nonlinearity(self.h, stage, _dp(psi_i.view(np.float64)), _dp(out.view(np.float64)))
        return out


def ref_setup(sp: SimulationParameters, raman: bool = True) -> dict:
    """Run the compiled reference constructor + initLawsonRK and return its products."""
    cp = CParams(sp, raman)
    M, nt = int(sp.n_modes), int(sp.nt)
    counts = np.zeros(4, dtype=np.uint32)
    scal = np.zeros(8)
    S = 11
    a = np.zeros(S * S)
    b = np.zeros(S)
    c = np.zeros(S)
    nb = len(sp.dispersion_coefficients[0])
    bt = np.zeros(M * nb)
    qt = np.zeros(max(cp.nnz, 1))
    pairs = np.zeros(max(cp.nnz, 1), dtype=np.uint32)
    psi = np.zeros((M, nt), dtype=np.complex128)
    L = np.zeros((M, nt), dtype=np.complex128)
    E1 = np.zeros((M, nt), dtype=np.complex128)
    E = np.zeros((S, M, nt), dtype=np.complex128)
    ref_lib().ref_probe_setup(ctypes.byref(cp.c), _up(counts), _dp(scal), _dp(a), _dp(b), _dp(c), _dp(bt), _dp(qt),
                              _up(pairs), _dp(psi.view(np.float64)), _dp(L.view(np.float64)),
                              _dp(E1.view(np.float64)), _dp(E.view(np.float64)))
    s = int(counts[0])
    return {
        "stages": s, "n_pairs": int(counts[1]), "default_steps": int(counts[2]),
        "t_final": scal[0], "delta_t": scal[1], "z_final": scal[2], "delta_z": scal[3], "z": scal[4],
        "conversion_factor": scal[5], "sign_gamma": scal[6], "T0": scal[7],
        "a": a[: s * s].reshape(s, s), "b": b[:s], "c": c[:s], "beta_tilde": bt.reshape(M, nb),
        "q_tilde": qt[: cp.nnz], "pairs": pairs[: int(counts[1])], "psi": psi, "L": L, "E1": E1, "E": E[:s],
    }
