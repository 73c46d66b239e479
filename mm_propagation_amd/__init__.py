"""MI355X (gfx950) Lawson-Runge-Kutta multimode fiber propagation.

Host mirror of the reference interface (SimulationParameters, Sparse3DArray,
MultimodePropagationGPU_RamanON/OFF) over the C ABI of libmmprop_hip.so
(include/mmprop.h). The propagation runs only in the HIP library: there is no
CPU fallback, and the solver classes raise if the library or a device is missing.
"""
from .problem import InputParameters, NonZeroValue, SimulationParameters, Sparse3DArray  # noqa: F401

__all__ = ["InputParameters", "NonZeroValue", "SimulationParameters", "Sparse3DArray"]
