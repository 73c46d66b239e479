"""Parity of the HIP path (through the C ABI) with the FP64 oracle on identical inputs.

Bar: per-mode relative L2 at the end of the run <= 1e-10 (the FP64 parity bar of the
plain solver); the short runs below observe ~1e-15, so they are held to 1e-12."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.problem import NonZeroValue
from mm_propagation_amd.solver import (MultimodePropagationGPU_RamanOFF, MultimodePropagationGPU_RamanON)

pytestmark = pytest.mark.gpu

TOL_SHORT = 1e-12
TOL_FULL = 1e-10


def run_pair(sp, steps, raman=True, decode=0, flags=0, n_threads=-1):
    require_device()
    cls = MultimodePropagationGPU_RamanON if raman else MultimodePropagationGPU_RamanOFF
    note(f"case M={sp.n_modes} NT={sp.nt} steps={steps} raman={raman} decode={decode} order={sp.method_order}")
    g = cls(sp, pair_decode=decode, flags=flags)
    g.computeLawsonRK(steps)
    got = g.getResult()
    g.close()
    o = ob.Oracle(sp, raman=raman, pair_decode=decode, n_threads=n_threads)
    o.compute_lawson_rk(steps)
    ref = o.get_result()
    return configs.rel_l2_per_mode(got, ref), got, ref


@pytest.mark.parametrize("nt", [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536])
def test_every_supported_length(nt):
    err, _, _ = run_pair(configs.two_mode_gaussian(nt=nt), 6)
    assert np.all(err <= TOL_SHORT), err


@pytest.mark.parametrize("order", [4, 6, 8])
@pytest.mark.parametrize("raman", [True, False])
@pytest.mark.parametrize("decode", [0, 1])
def test_tableaux_raman_and_pair_decode(order, raman, decode):
    sp = configs.two_mode_gaussian(nt=512, method_order=order)
    err, _, _ = run_pair(sp, 25, raman=raman, decode=decode)
    assert np.all(err <= TOL_SHORT), err


@pytest.mark.parametrize("M", [1, 3, 6, 10])
def test_grin_modes(M):
    sp = configs.grin_fiber(M=M, nt=1024, symmetric=True)
    err, _, _ = run_pair(sp, 30, decode=1)
    assert np.all(err <= TOL_SHORT), err


def test_grin_k_le_l_rule_reference_decode_short():
    """The literal k <= l GRIN rule with the reference-CPU decode (diverges later in z; 20 steps stay finite)."""
    sp = configs.grin_fiber(M=6, nt=2048)
    err, got, _ = run_pair(sp, 20, decode=0)
    assert np.isfinite(got).all()
    assert np.all(err <= TOL_SHORT), err


def test_graph_replays_match_eager():
    """HIP-graph replay (16-step graphs) against one-by-one launches, several replays with
    eager work and synchronisations in between; results must agree bit for bit."""
    require_device()
    sp = configs.grin_fiber(M=4, nt=2048, symmetric=True)
    note("case graph vs eager, M=4 NT=2048")
    a = MultimodePropagationGPU_RamanON(sp, pair_decode=1)
    b = MultimodePropagationGPU_RamanON(sp, pair_decode=1, flags=_capi.MMP_FLAG_NO_GRAPH)
    a.computeLawsonRK(0)
    b.computeLawsonRK(0)
    for rep in range(4):
        a.advance(32)  # two graph replays
        b.advance(32)
        ea = a.getResult()  # eager kernels + synchronisation between replays
        eb = b.getResult()
        assert np.array_equal(ea, eb), rep
    a.close()
    b.close()


def test_two_mode_gaussian_full_length():
    """The 2-mode Gaussian workload (M=2, NT=4096, 5000 RK4 steps to z = L, reference-CPU pair decode)."""
    sp = configs.two_mode_gaussian()
    err, got, ref = run_pair(sp, sp.n_steps, n_threads=8)
    assert np.all(err <= TOL_FULL), err


def break_parity(sp):
    """Adds Q^(0)_{1,0,0} (k + l - m - p odd) so that no slot keeps a single parity of p + k."""
    sp.coupling_coefficients[0].addNonZeroValue(NonZeroValue((1, 0, 0), 0.25 * sp.coupling_coefficients[0]
                                                            .getNonZeroValue(0).value))
    return sp


@pytest.mark.parametrize("M,nt,broken", [(2, 4096, False), (6, 16384, False), (10, 8192, False), (12, 4096, False),
                                         (13, 4096, False), (14, 4096, True), (15, 8192, False), (16, 4096, False)])
def test_fused_column_passes_match_separate_passes(M, nt, broken):
    """The column-stage kernels (4 passes per stage) against the separate passes (7 per stage),
    and both against the oracle; 10 modes and more run the Raman lines in several LDS chunks;
    `broken` couplings take the full M x M coupling blocks instead of the parity-compact ones."""
    require_device()
    sp = configs.two_mode_gaussian(nt=nt) if M == 2 else configs.grin_fiber(M=M, nt=nt, symmetric=True)
    if broken:
        sp = break_parity(sp)
    dec = 0 if M == 2 else 1
    note(f"case fused vs separate M={M} NT={nt}")
    a = MultimodePropagationGPU_RamanON(sp, pair_decode=dec)
    b = MultimodePropagationGPU_RamanON(sp, pair_decode=dec, flags=_capi.MMP_FLAG_SEPARATE_COLUMNS)
    a.computeLawsonRK(5)
    b.computeLawsonRK(5)
    ra, rb = a.getResult(), b.getResult()
    a.close()
    b.close()
    assert np.all(configs.rel_l2_per_mode(ra, rb) <= 1e-14)
    o = ob.Oracle(sp, pair_decode=dec)
    o.compute_lawson_rk(5)
    assert np.all(configs.rel_l2_per_mode(ra, o.get_result()) <= TOL_SHORT)


@pytest.mark.parametrize("order", [4, 6, 8])
@pytest.mark.parametrize("decode", [0, 1])
def test_column_stage_kernels_tableaux_and_decode(order, decode):
    """Default path at NT = 4096 (column-stage kernels) and the separate passes, both vs the oracle."""
    sp = configs.two_mode_gaussian(nt=4096, method_order=order)
    err, _, _ = run_pair(sp, 25, decode=decode)
    assert np.all(err <= TOL_SHORT), err
    err, _, _ = run_pair(sp, 25, decode=decode, flags=_capi.MMP_FLAG_SEPARATE_COLUMNS)
    assert np.all(err <= TOL_SHORT), err


def test_column_stage_kernels_grin_k_le_l_reference_decode():
    sp = configs.grin_fiber(M=6, nt=16384)
    err, _, _ = run_pair(sp, 20, decode=0)
    assert np.all(err <= TOL_SHORT), err
    err, _, _ = run_pair(sp, 20, decode=0, flags=_capi.MMP_FLAG_SEPARATE_COLUMNS)
    assert np.all(err <= TOL_SHORT), err
