"""The library's host-side preparation (C ABI mmp_plan / mmp_plan_arrays, no device):
validation statuses, scalars, and the prepared psi0, L and Lawson tables, bit for bit
against the reference's own setup code (oracle/_ref) and the oracle."""
import math

import numpy as np
import pytest

import oracle_binding as ob
from mm_propagation_amd import _capi, configs, plan
from test_oracle_setup import CASES


@pytest.mark.parametrize("name", sorted(CASES))
def test_prepared_arrays_match_oracle_bitwise(name):
    sp = CASES[name]()
    for decode in (0, 1):
        got = plan.plan_arrays(sp, pair_decode=decode)
        o = ob.Oracle(sp, pair_decode=decode)
        o.init_operators()
        assert np.array_equal(got["psi0"], o.get_state()[0])
        L = o.linear_operator()
        assert np.array_equal(got["imL"], L.imag)
        assert np.array_equal(got["e1"], o.e1())
        for i in range(o.info.stages):
            assert np.array_equal(got["e"][i], o.e(i))
        a, b = o.pair_table()
        assert np.array_equal(got["pair_a"], a) and np.array_equal(got["pair_b"], b)
        info = got["info"]
        for k in ("t_final", "delta_t", "z_final", "delta_z", "conversion_factor", "sign_gamma"):
            assert info[k] == getattr(o.info, k), k
        assert info["default_steps"] == o.default_steps()


@pytest.mark.skipif(not ob.ref_available(), reason="reference setup probe not built (make -C oracle ref)")
@pytest.mark.parametrize("name", sorted(CASES))
def test_prepared_arrays_match_reference_setup_bitwise(name):
    sp = CASES[name]()
    got = plan.plan_arrays(sp)
    r = ob.ref_setup(sp)
    assert np.array_equal(got["psi0"], r["psi"])
    assert np.array_equal(got["imL"], r["L"].imag)
    assert np.array_equal(got["e1"], r["E1"])
    assert np.array_equal(got["e"], r["E"])


@pytest.mark.parametrize("nt,n1", [(64, 8), (128, 8), (256, 16), (1024, 32), (2048, 32), (4096, 64), (8192, 64),
                                   (16384, 64), (32768, 64), (65536, 64)])
def test_four_step_split(nt, n1):
    """Columns of at most 64 samples while rows stay within 1024 samples."""
    i = plan.plan(configs.two_mode_gaussian(nt=nt))
    assert (i["n1"], i["n2"]) == (n1, nt // n1)


def test_plan_sizes_and_memory():
    sp = configs.grin_fiber(M=6, nt=16384, symmetric=True)
    i = plan.plan(sp, pair_decode=1)
    assert (i["n_modes"], i["nt"], i["stages"], i["n_pairs"], i["nnz"]) == (6, 16384, 4, 36, 648)
    assert (i["n1"], i["n2"]) == (64, 256)
    # 21 Raman lines: 15 cross pairs + 3 lines of packed self pairs
    assert i["n_conv_lines"] == 18
    assert 10e6 < i["device_bytes"] < 40e6


FP64, MIXED = _capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED


@pytest.mark.parametrize("M,nt,precision,expect", [
    (6, 16384, FP64, (4, 4)),    # 64 blocks of 4 columns; 18 lines in 4 groups: 256 workgroups
    (6, 16384, MIXED, (4, 4)),
    (10, 65536, MIXED, (4, 2)),  # 256 blocks; 50 lines in 2 groups: 512 workgroups
    (10, 65536, FP64, (2, 1)),   # 4 FP64 columns of 10 fields need 77.8 KB of LDS; 2 need 57.3 KB
    (15, 65536, FP64, (1, 1)),   # 2 columns need 67.6 KB; 1024 one-column blocks
    (2, 4096, FP64, (4, 1)),     # 2 lines: a second group would hold fewer than 4
])
def test_pair_pass_layout(M, nt, precision, expect):
    """The column-stage pair pass of a one-fiber solver: widest column block whose LDS fits in
    64 KB, then line groups of at least 4 lines until 512 workgroups (mmp_plan, host only)."""
    i = plan.plan(configs.grin_fiber(M=M, nt=nt, symmetric=True), pair_decode=1, precision=precision)
    assert (i["pair_columns"], i["pair_groups"]) == expect
    assert i["gather_hi_occupancy"] == 0  # decided with the fiber count, by the solver


def expect_status(sp, status, raman=True):
    with pytest.raises(_capi.MmpError) as ei:
        plan.plan(sp, raman=raman)
    assert ei.value.status == status, str(ei.value)


def test_validation_statuses():
    base = configs.two_mode_gaussian(nt=256)
    sp = configs.two_mode_gaussian(nt=256)
    sp.nt = 100
    sp.initial_fields = np.zeros((2, 100), complex)
    sp.raman_response = np.zeros(100)
    expect_status(sp, 2)  # NT not a power of two: unsupported
    sp = configs.two_mode_gaussian(nt=32)
    expect_status(sp, 2)  # below the smallest supported length
    sp = configs.two_mode_gaussian(nt=256)
    sp.dispersion_coefficients = [[0, 0], [0, 0]]
    expect_status(sp, 1)  # fewer than 3 orders
    sp = configs.two_mode_gaussian(nt=256)
    sp.dispersion_coefficients = [[0, 0, 0], [0, 0, 0]]
    expect_status(sp, 1)  # max |beta2| = 0
    sp = configs.two_mode_gaussian(nt=256)
    for q in sp.coupling_coefficients:
        for i in range(len(q)):
            q.resetValue(i, 0.0)
    expect_status(sp, 1)  # max |Q| = 0
    sp = configs.two_mode_gaussian(nt=256)
    sp.coupling_coefficients[0].addNonZeroValue(configs.NonZeroValue((0, 2, 0), 1.0))
    expect_status(sp, 1)  # index out of range
    sp = configs.two_mode_gaussian(nt=256)
    sp.initial_fields = sp.initial_fields.copy()
    sp.initial_fields[1, 3] = np.nan
    expect_status(sp, 6)  # non-finite input
    sp = configs.two_mode_gaussian(nt=256)
    sp.n_steps = 0
    expect_status(sp, 1)
    # Raman off needs no response
    sp = configs.two_mode_gaussian(nt=256)
    sp.raman_response = None
    assert plan.plan(sp, raman=False)["n_conv_lines"] == 0
    assert plan.plan(base)["n_conv_lines"] == 2


def test_mixed_precision_plan_holds_no_lawson_tables():
    """Mixed precision forms the Lawson factors from Im L and z: its footprint lacks the FP64
    form's three RK4 tables, their restart copies and E1 (7 x 16 B per sample), on top of the
    FP32 work arrays' saving."""
    M, nt = 6, 16384
    sp = configs.grin_fiber(M=M, nt=nt, symmetric=True)
    fp64 = plan.plan(sp, pair_decode=1)["device_bytes"]
    mixed = plan.plan(sp, pair_decode=1, precision=MIXED)["device_bytes"]
    assert fp64 - mixed >= 7 * 16 * M * nt
