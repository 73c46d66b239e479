"""Reference semantics of the device solver: restart on every computeLawsonRK(nz),
z by repeated addition, the default step count floor(z_final / delta_z), resumable stepping, state
round trips, energy diagnostics and error statuses."""
import ctypes
import math

import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs, plan
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON, make_solver

pytestmark = pytest.mark.gpu


def test_restart_semantics_match_oracle():
    require_device()
    sp = configs.two_mode_gaussian(nt=1024)
    note("case restart semantics M=2 NT=1024")
    g = MultimodePropagationGPU_RamanON(sp)
    o = ob.Oracle(sp, pair_decode=1)  # the GPU classes decode pairs as the reference GPU class
    for nz in (7, 5, 11):
        g.computeLawsonRK(nz)
        o.compute_lawson_rk(nz)
    err = configs.rel_l2_per_mode(g.getResult(), o.get_result())
    assert np.all(err < 1e-12), err
    psi_g, z_g = g.get_state()
    psi_o, z_o = o.get_state()
    assert z_g == z_o  # bitwise: both sum h eleven times
    assert np.all(configs.rel_l2_per_mode(psi_g, psi_o) < 1e-12)
    g.close()


def test_resumable_advance_equals_single_run():
    require_device()
    sp = configs.grin_fiber(M=3, nt=1024, symmetric=True)
    note("case advance split M=3 NT=1024")
    a = MultimodePropagationGPU_RamanON(sp, pair_decode=1)
    a.computeLawsonRK(10)
    a.advance(23)
    b = MultimodePropagationGPU_RamanON(sp, pair_decode=1)
    b.computeLawsonRK(33)
    assert np.array_equal(a.getResult(), b.getResult())
    assert a.info()["z"] == b.info()["z"]
    a.close()
    b.close()


def test_default_step_count_and_info():
    require_device()
    sp = configs.two_mode_gaussian(nt=256, n_steps=40, fiber_length=0.004)  # same h as the 0.5 m run
    note("case default steps")
    g = MultimodePropagationGPU_RamanON(sp)
    i = g.info()
    assert i["default_steps"] == math.floor(i["z_final"] / i["delta_z"])
    g.computeLawsonRK()
    assert g.info()["steps_taken"] == i["default_steps"]
    o = ob.Oracle(sp, pair_decode=1)  # the GPU classes decode pairs as the reference GPU class
    o.compute_lawson_rk(o.default_steps())
    assert np.all(configs.rel_l2_per_mode(g.getResult(), o.get_result()) < 1e-12)
    g.close()


def test_state_round_trip_and_energy():
    require_device()
    sp = configs.grin_fiber(M=4, nt=2048, symmetric=True)
    note("case state round trip and energy")
    g = MultimodePropagationGPU_RamanON(sp, pair_decode=1)
    psi0, z0 = g.get_state()
    assert z0 == 0.0
    conv = g.info()["conversion_factor"]
    assert np.allclose(psi0, sp.initial_fields * conv, rtol=1e-14, atol=1e-14 * np.abs(psi0).max())
    g.computeLawsonRK(12)
    res = g.getResult()
    e = g.energy()
    assert np.allclose(e, configs.energy_per_mode(res, sp.time_window), rtol=1e-13)
    g.set_state(psi0)
    assert g.info()["z"] == 0.0
    back, _ = g.get_state()
    assert np.allclose(back, psi0, rtol=0, atol=1e-14 * np.abs(psi0).max())
    g.close()


def test_nonfinite_field_is_reported():
    """A step of h = z_final / 3 drives the two-mode workload to overflow: getResult must
    report MMP_ERR_NONFINITE, as the oracle's field is non-finite too."""
    require_device()
    sp = configs.two_mode_gaussian(nt=256, n_steps=3)
    note("case non-finite detection")
    g = MultimodePropagationGPU_RamanON(sp)
    g.computeLawsonRK(3)
    with pytest.raises(_capi.MmpError) as ei:
        g.getResult()
    assert ei.value.status == 6
    g.close()


@pytest.mark.parametrize("mutate,status", [
    (lambda sp: setattr(sp, "nt", 100) or setattr(sp, "initial_fields", np.zeros((2, 100), complex)) or
     setattr(sp, "raman_response", np.zeros(100)), 2),
    (lambda sp: setattr(sp, "dispersion_coefficients", [[0, 0], [0, 0]]), 1),
    (lambda sp: sp.coupling_coefficients[0].resetValue(0, 0.0) or sp.coupling_coefficients[0].resetValue(1, 0.0) or
     sp.coupling_coefficients[0].resetValue(2, 0.0) or [sp.coupling_coefficients[1].resetValue(i, 0.0) for i in range(3)],
     1),
])
def test_invalid_inputs_rejected_on_device_path(mutate, status):
    require_device()
    sp = configs.two_mode_gaussian(nt=256)
    mutate(sp)
    with pytest.raises(_capi.MmpError) as ei:
        MultimodePropagationGPU_RamanON(sp)
    assert ei.value.status == status


@pytest.mark.parametrize("precision,flags", [(0, 0), (1, 0), (0, _capi.MMP_FLAG_SEPARATE_COLUMNS)])
def test_planned_bytes_equal_allocated_bytes(precision, flags):
    """mmp_plan's device_bytes (host only) is what a solver with the same options allocates."""
    require_device()
    sp = configs.grin_fiber(M=6, nt=4096, symmetric=True)
    note(f"case planned bytes precision={precision} flags={flags}")
    planned = plan.plan(sp, pair_decode=1, precision=precision, flags=flags)["device_bytes"]
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=precision, flags=flags)
    got = s.info()["device_bytes"]
    s.close()
    assert planned == got


@pytest.mark.parametrize("resumable", [False, True])
def test_snapshots_match_oracle(resumable):
    """n_save snapshots: the reference restart semantics of the MEX gateway (each chunk a fresh
    computeLawsonRK) or one resumable solve; both against the oracle driven the same way,
    with the remainder steps in the last chunk and Z accumulated in metres."""
    require_device()
    n_steps, n_save = 50, 4  # chunks of 12, 12, 12, 14
    sp = configs.two_mode_gaussian(nt=1024, n_steps=n_steps, fiber_length=0.005)  # h = 1e-4 m
    note(f"case snapshots resumable={resumable}")
    s = MultimodePropagationGPU_RamanON(sp)
    fields, z = s.propagate_snapshots(n_steps, n_save, resumable=resumable)
    s.close()
    o = ob.Oracle(sp, pair_decode=1)  # the GPU classes decode pairs as the reference GPU class
    Z = 0.0
    for i, n in enumerate([12, 12, 12, 14]):
        if resumable and i > 0:
            o.advance(n)
        else:
            o.compute_lawson_rk(n)
        Z += n * (sp.fiber_length / n_steps)
        assert z[i] == Z
        assert np.all(configs.rel_l2_per_mode(fields[i], o.get_result()) <= 1e-12)
    o.close()
    if resumable:
        one = MultimodePropagationGPU_RamanON(sp)
        one.computeLawsonRK(n_steps)
        assert np.all(configs.rel_l2_per_mode(fields[-1], one.getResult()) <= 1e-13)
        one.close()


def test_snapshot_arguments_are_checked():
    require_device()
    s = MultimodePropagationGPU_RamanON(configs.two_mode_gaussian(nt=256))
    for n_steps, n_save, mode in ((10, 0, 0), (10, 11, 0), (10, 2, 7)):
        with pytest.raises(_capi.MmpError) as ei:
            _capi.check(s.lib.mmp_propagate_snapshots(s.h, n_steps, n_save, mode, None, None), "snapshots")
        assert ei.value.status == _capi.MMP_ERR_INVALID_ARGUMENT
    s.close()
