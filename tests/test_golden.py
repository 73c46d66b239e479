"""Checks against the committed fixtures of tests/golden/ (written by
tests/golden/make_golden.py): the reference's own setup products, and the oracle's
two-mode Gaussian field at z = L."""
import glob
import os
import sys

import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import configs, plan

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

SETUP = make_golden.setup_cases()
TWO_MODE = os.path.join(GOLDEN, "oracle_two_mode_gaussian_zL.npz")


def test_every_setup_case_has_a_fixture():
    have = {os.path.basename(p)[len("reference_setup_"):-4] for p in glob.glob(os.path.join(GOLDEN, "reference_setup_*.npz"))}
    assert have == set(SETUP)


@pytest.mark.parametrize("name", sorted(SETUP))
def test_oracle_setup_equals_reference_vectors(name):
    """The oracle's constructor products equal the reference's, bit for bit."""
    make, raman = SETUP[name]
    sp = make()
    r = np.load(os.path.join(GOLDEN, f"reference_setup_{name}.npz"))
    o = ob.Oracle(sp, raman=raman)
    o.init_operators()
    a, b, c = o.tableau()
    assert int(r["stages"]) == o.info.stages
    assert np.array_equal(a, r["a"]) and np.array_equal(b, r["b"]) and np.array_equal(c, r["c"])
    assert np.array_equal(o.beta_tilde(), r["beta_tilde"])
    assert np.array_equal(o.q_tilde(), r["q_tilde"])
    assert np.array_equal(o.unique_pairs(), r["pairs"])
    assert np.array_equal(o.get_state()[0], r["psi"])
    assert np.array_equal(o.linear_operator(), r["L"])
    assert np.array_equal(o.e1(), r["E1"])
    for i in range(o.info.stages):
        assert np.array_equal(o.e(i), r["E"][i])
    for k in ("t_final", "delta_t", "z_final", "delta_z", "conversion_factor", "sign_gamma"):
        assert getattr(o.info, k) == float(r[k]), k
    assert o.default_steps() == int(r["default_steps"])
    o.close()


@pytest.mark.parametrize("name", sorted(SETUP))
def test_host_plan_equals_reference_vectors(name):
    """The product's host preparation (mmp_plan_arrays) equals the reference's, bit for bit."""
    make, raman = SETUP[name]
    sp = make()
    r = np.load(os.path.join(GOLDEN, f"reference_setup_{name}.npz"))
    got = plan.plan_arrays(sp, raman=raman)
    assert np.array_equal(got["psi0"], r["psi"])
    assert np.array_equal(got["imL"], r["L"].imag) and not np.any(r["L"].real)
    assert np.array_equal(got["e1"], r["E1"])
    assert np.array_equal(got["e"], r["E"])
    info = got["info"]
    for k in ("t_final", "delta_t", "z_final", "delta_z", "conversion_factor", "sign_gamma"):
        assert info[k] == float(r[k]), k
    assert info["default_steps"] == int(r["default_steps"])


@pytest.mark.slow
def test_oracle_reproduces_two_mode_fixture():
    """Regression of the oracle itself over the full 5000 steps of the two-mode workload."""
    g = np.load(TWO_MODE)
    sp = configs.two_mode_gaussian()
    o = ob.Oracle(sp, n_threads=2)
    o.compute_lawson_rk(int(g["steps"]))
    assert o.get_state()[1] == float(g["z"])
    assert np.all(configs.rel_l2_per_mode(o.get_result(), g["fields"]) <= 1e-14)
    o.close()


@pytest.mark.gpu
def test_device_two_mode_full_length_against_fixture():
    """Device field at z = L (5000 RK4 steps, Raman on, reference decode) within 1e-10 per mode."""
    require_device()
    from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON
    g = np.load(TWO_MODE)
    sp = configs.two_mode_gaussian()
    note("case two-mode fixture, 5000 steps")
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=0)  # the fixture uses the reference CPU decode
    s.computeLawsonRK(int(g["steps"]))
    fields = s.getResult()
    z = s.info()["z"]
    s.close()
    assert z == float(g["z"])
    err = configs.rel_l2_per_mode(fields, g["fields"])
    assert np.all(err <= 1e-10), err
