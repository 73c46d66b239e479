"""Device diagnostics of the propagated field: the FP64 photon-number reduction against a numpy
reduction of getResult's field, and its argument checks."""
import numpy as np
import pytest

from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.batch import MultimodePropagationBatch
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu

HBAR = 1.054571817e-34  # J s
C_LIGHT = 299792458.0   # m/s


def photon_number_numpy(fields: np.ndarray, time_window: float, wavelength: float) -> np.ndarray:
    """N_p = sum_k E_p,k / (hbar omega_k) from the physical field A_p (sqrt(W)) on NT samples:
    E_p,k = (dt / NT) |FFT(A_p)_k|^2 (pJ), omega_k = omega0 - 2 pi k / T_w (rad/ps) with k in
    (-NT/2, NT/2] in numpy's bin order."""
    nt = fields.shape[1]
    dt = time_window / nt
    k = np.arange(nt)
    k = np.where(k <= nt // 2, k, k - nt).astype(np.float64)
    omega0 = 2 * np.pi * C_LIGHT / wavelength * 1e-12
    omega = omega0 - 2 * np.pi * k / time_window
    spec = np.abs(np.fft.fft(fields, axis=1)) ** 2 * (dt / nt)  # pJ per bin
    return (spec * 1e-12 / (HBAR * omega * 1e12)).sum(axis=1)


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
def test_photon_number_matches_numpy_reduction(precision):
    sp = configs.grin_fiber(M=6, nt=4096, symmetric=True)
    require_device()
    note(f"case photon number M=6 NT=4096 precision={precision}")
    s = MultimodePropagationGPU_RamanON(sp, precision=precision)
    lam = 1.03e-6
    n0 = s.photon_number(lam)
    ref0 = photon_number_numpy(sp.initial_fields, sp.time_window, lam)
    assert np.allclose(n0, ref0, rtol=1e-12, atol=0), (n0, ref0)
    s.computeLawsonRK(20)
    got = s.photon_number(lam)
    ref = photon_number_numpy(s.getResult(), sp.time_window, lam)
    assert np.allclose(got, ref, rtol=1e-12, atol=0), (got, ref)
    # a quasi-monochromatic check of the scale: N = E / (hbar omega0) to the spectral width
    e = s.energy() * 1e-12
    omega0 = 2 * np.pi * C_LIGHT / lam
    assert np.allclose(got, e / (HBAR * omega0), rtol=5e-2)
    with pytest.raises(_capi.MmpError) as ei:
        s.photon_number(-1.0)
    assert ei.value.status == _capi.MMP_ERR_INVALID_ARGUMENT
    # a carrier below the grid's half bandwidth: some bins would have omega <= 0
    with pytest.raises(_capi.MmpError) as ei:
        s.photon_number(1e-2)
    assert ei.value.status == _capi.MMP_ERR_INVALID_ARGUMENT
    s.close()


def test_fiber_photon_number_in_a_batch():
    sps = [configs.grin_fiber(M=3, nt=4096, symmetric=True, seed=i) for i in range(3)]
    require_device()
    note("case batch photon number M=3 NT=4096 x 3")
    b = MultimodePropagationBatch(sps)
    b.computeLawsonRK(16)
    for f, sp in enumerate(sps):
        got = b.photon_number(f, 1.55e-6)
        ref = photon_number_numpy(b.result(f), sp.time_window, 1.55e-6)
        assert np.allclose(got, ref, rtol=1e-12, atol=0), (f, got, ref)
    b.close()
