"""Lawson factors formed from the phase (the mixed-precision stage heads) against the tables the
reference advances (src/MultimodePropagation_CPU.cpp:37-44: E_i <- E_i * E1 once per step),
on the host with the library's own prepared arrays (mmp_plan_arrays)."""
import numpy as np

from mm_propagation_amd import configs, plan

FP32_UNIT = 2.0 ** -24


def test_phase_factors_stay_within_rounding_of_advanced_tables():
    """exp(i Im L (n + c_i) h), the mixed-precision heads' factor after n steps, against
    E_i(0) * E1^n after n steps of a 20000-step GRIN solve: the two differ only by the tables'
    rounding, which grows linearly with the steps and stays far below the FP32 stages' unit
    roundoff. (A phase from z summed step by step in FP64 would sit 8.6e-10 away at the end.)"""
    steps = 20000
    sp = configs.grin_fiber(M=3, nt=4096, n_steps=steps, symmetric=True)
    arr = plan.plan_arrays(sp, pair_decode=1)
    h = arr["info"]["delta_z"]
    im_l = arr["imL"]
    tables = arr["e"].copy()  # [stage][M][NT], E_i at z = 0
    e1 = arr["e1"]
    c = np.array([0.0, 0.5, 0.5, 1.0])  # RK4
    # the prepared tables are the phases at z = 0
    assert np.abs(tables - np.exp(1j * im_l[None] * (c[:, None, None] * h))).max() < 1e-15
    worst = []
    for n in range(1, steps + 1):
        tables *= e1
        if n % 5000 == 0:
            direct = np.exp(1j * im_l[None] * ((n + c[:, None, None]) * h))
            worst.append(float(np.abs(tables - direct).max()))
    print("max |E_table - E_phase| every 5000 steps:", worst)
    assert worst[-1] <= 1e-11
    assert worst[-1] < FP32_UNIT / 1000
    # growth with the steps (a wrong sign, c or step count would show as an O(1) difference)
    assert worst[0] < worst[-1]
