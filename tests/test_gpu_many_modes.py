"""More than 16 modes: the separate-pass step with the sparse LDS gather (nonzeros sorted by
slot, slot chunks of up to 108 slots, lane-broadcast coefficients), against the FP64 oracle.
FP64 at the 1e-12 bar of short runs; mixed precision at 1e-6 or the dynamics' amplification of
an FP64 rounding-level perturbation scaled to FP32 (tests/test_gpu_mixed.py)."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanOFF, MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu

FP64_SHORT = 1e-12


def oracle(sp, steps, raman, variant=0):
    o = ob.Oracle(sp, raman=raman, pair_decode=1, fft_variant=variant, n_threads=16)
    o.compute_lawson_rk(steps)
    r = o.get_result()
    o.close()
    return r


@pytest.mark.parametrize("M,nt,steps", [(20, 256, 3), (32, 128, 2)])
@pytest.mark.parametrize("raman", [True, False])
@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
def test_many_modes_against_oracle(M, nt, steps, raman, precision):
    require_device()
    sp = configs.grin_fiber(M=M, nt=nt, symmetric=True, time_window=5.0, power_range=(1.0, 100.0))
    note(f"case many modes M={M} NT={nt} steps={steps} raman={raman} precision={precision}")
    cls = MultimodePropagationGPU_RamanON if raman else MultimodePropagationGPU_RamanOFF
    s = cls(sp, precision=precision)
    info = s.info()
    assert info["gather_waves"] == 0  # the separate passes
    s.computeLawsonRK(steps)
    got = s.getResult()
    s.close()
    ref = oracle(sp, steps, raman)
    err = configs.rel_l2_per_mode(got, ref)
    print(f"M={M} rel-L2 max {err.max():.3e}", flush=True)
    if precision == _capi.MMP_PRECISION_FP64:
        assert np.all(err <= FP64_SHORT), err
    else:
        s64 = float(configs.rel_l2_per_mode(oracle(sp, steps, raman, variant=1), ref).max())
        assert np.all(err <= max(1e-6, 4.0 * 2.0 ** 29 * s64)), (err, s64)
