"""Host-side mirror of the reference's input types.

`SimulationParameters` follows include/MultimodePropagation.hpp:16-33 (the
README calls it InputParameters, README.md:71-86); `Sparse3DArray` and
`NonZeroValue` follow include/Sparse3DArray.hpp:7-30 (a COO list kept in
insertion order, src/Sparse3DArray.cpp:16-19). Fields use the reference's names
and units. Complex containers are numpy arrays of shape (M, NT), complex128.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _capi


@dataclass
class NonZeroValue:
    """One coupling coefficient Q^(p)_{k,l,m} (include/Sparse3DArray.hpp:7-11)."""

    index: tuple
    value: float


class Sparse3DArray:
    """COO list of (k, l, m, value) in insertion order (include/Sparse3DArray.hpp:14-30)."""

    def __init__(self, entries: Optional[Sequence] = None):
        self._q: List[NonZeroValue] = []
        for e in entries or ():
            if isinstance(e, NonZeroValue):
                self.addNonZeroValue(e)
            else:
                k, l, m, v = e
                self.addNonZeroValue(NonZeroValue((int(k), int(l), int(m)), float(v)))

    def addNonZeroValue(self, nzv: NonZeroValue) -> None:
        self._q.append(NonZeroValue(tuple(int(i) for i in nzv.index), float(nzv.value)))

    def getNonZeroValue(self, i: int) -> NonZeroValue:
        return self._q[i]

    def getNumberOfNonZeroValues(self) -> int:
        return len(self._q)

    def resetValue(self, i: int, value: float) -> None:
        self._q[i].value = float(value)

    def __len__(self) -> int:
        return len(self._q)

    def __iter__(self):
        return iter(self._q)


@dataclass
class SimulationParameters:
    """Input record of the solver (include/MultimodePropagation.hpp:16-33)."""

    n_modes: int = 0
    fiber_length: float = 0.0
    dispersion_coefficients: List[List[float]] = field(default_factory=list)
    coupling_coefficients: List[Sparse3DArray] = field(default_factory=list)
    nt: int = 0
    time_window: float = 0.0
    pulse_width: float = 0.0
    initial_fields: Optional[np.ndarray] = None  # (M, NT) complex128
    n_steps: int = 0
    method_order: int = 4
    nonlinearity_const: float = 0.0
    raman_proportion: float = 0.0
    raman_response: Optional[np.ndarray] = None  # (NT,) float64


InputParameters = SimulationParameters  # README.md:71 name


class CParams:
    """An MmpParams view of a SimulationParameters; owns the flattened arrays it points to."""

    def __init__(self, sp: SimulationParameters, raman: bool):
        M = int(sp.n_modes)
        if len(sp.dispersion_coefficients) != M:
            raise ValueError("dispersion_coefficients must have n_modes rows")
        n_beta = len(sp.dispersion_coefficients[0]) if M else 0
        self.beta = np.ascontiguousarray(np.array(sp.dispersion_coefficients, dtype=np.float64).reshape(M, n_beta))
        if len(sp.coupling_coefficients) != M:
            raise ValueError("coupling_coefficients must have n_modes entries")
        offs = [0]
        klm: List[int] = []
        vals: List[float] = []
        for q in sp.coupling_coefficients:
            for nzv in q:
                klm.extend(nzv.index)
                vals.append(nzv.value)
            offs.append(len(vals))
        self.offsets = np.array(offs, dtype=np.uint32)
        self.klm = np.array(klm if klm else [0, 0, 0], dtype=np.uint32)
        self.values = np.array(vals if vals else [0.0], dtype=np.float64)
        fields = np.asarray(sp.initial_fields, dtype=np.complex128)
        if fields.shape != (M, int(sp.nt)):
            raise ValueError(f"initial_fields must have shape ({M}, {sp.nt}), got {fields.shape}")
        self.fields = np.ascontiguousarray(fields).view(np.float64)
        self.raman_response = None
        if raman:
            if sp.raman_response is None:
                raise ValueError("raman_response is required with Raman on")
            self.raman_response = np.ascontiguousarray(np.asarray(sp.raman_response, dtype=np.float64))
            if self.raman_response.shape != (int(sp.nt),):
                raise ValueError("raman_response must have NT entries")
        dp = ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint32)
        self.c = _capi.MmpParams(
            n_modes=M,
            fiber_length=float(sp.fiber_length),
            n_beta=n_beta,
            dispersion_coefficients=self.beta.ctypes.data_as(dp),
            coupling_offsets=self.offsets.ctypes.data_as(up),
            coupling_klm=self.klm.ctypes.data_as(up),
            coupling_values=self.values.ctypes.data_as(dp),
            nt=int(sp.nt),
            time_window=float(sp.time_window),
            pulse_width=float(sp.pulse_width),
            initial_fields=self.fields.ctypes.data_as(dp),
            n_steps=int(sp.n_steps),
            method_order=int(sp.method_order),
            nonlinearity_const=float(sp.nonlinearity_const),
            raman_proportion=float(sp.raman_proportion),
            raman_response=self.raman_response.ctypes.data_as(dp) if self.raman_response is not None else dp(),
            raman=1 if raman else 0,
        )

    @property
    def nnz(self) -> int:
        return int(self.offsets[-1])
