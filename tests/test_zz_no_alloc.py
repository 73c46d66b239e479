"""mmp_solve_static (the no-allocation entry point) against the solver it mirrors
(mmp_solver_create + mmp_compute_lawson_rk + mmp_get_result), bit for bit, on two calls in a
row, the second with a smaller problem (the static scratch keeps the first call's contents)."""
import numpy as np
import pytest

from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.solver import (Context, MultimodePropagationGPU_RamanOFF, MultimodePropagationGPU_RamanON,
                                       solve_static)

pytestmark = pytest.mark.gpu


def solve_with_solver(sp, nz, raman, precision, ctx):
    cls = MultimodePropagationGPU_RamanON if raman else MultimodePropagationGPU_RamanOFF
    s = cls(sp, precision=precision, ctx=ctx)
    s.computeLawsonRK(nz)
    r = s.getResult()
    s.close()
    return r


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
def test_static_solve_equals_the_solver(precision):
    require_device()
    ctx = Context(0)
    cases = [(configs.grin_fiber(M=6, nt=4096, symmetric=True), 20, True),
             (configs.two_mode_gaussian(nt=1024), 12, False)]  # the second call is smaller
    for sp, nz, raman in cases:
        note(f"case static solve M={sp.n_modes} NT={sp.nt} steps={nz} raman={raman} precision={precision}")
        got = solve_static(sp, nz, raman=raman, precision=precision, ctx=ctx)
        ref = solve_with_solver(sp, nz, raman, precision, ctx)
        assert np.array_equal(got, ref), configs.rel_l2_per_mode(got, ref)
    ctx.close()


def test_static_solve_refuses_what_does_not_fit():
    require_device()
    sp = configs.grin_fiber(M=10, nt=65536, symmetric=True)  # hundreds of MB in FP64
    with pytest.raises(_capi.MmpError) as ei:
        solve_static(sp, 1)
    assert ei.value.status == _capi.MMP_ERR_UNSUPPORTED
    with pytest.raises(_capi.MmpError) as ei:
        solve_static(configs.two_mode_gaussian(nt=256), 1, flags=_capi.MMP_FLAG_AUTOTUNE)
    assert ei.value.status == _capi.MMP_ERR_INVALID_ARGUMENT
