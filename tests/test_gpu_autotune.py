"""MMP_FLAG_AUTOTUNE: the tuner times the column-stage variants (gather wavefronts, lines per
chunk and instance; pair-pass columns per workgroup and line groups; the line order into the Raman
row pass) on the solver's own data, keeps the fastest, and leaves the solver exactly as created
(z = 0, the initial field, no fiber parked); a later solver of the same structure takes the cached
choice. Tuned results stay within rounding of the oracle (only the gather variants change an
order of summation, over the pair slots)."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.batch import MultimodePropagationBatch
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu


def test_tuned_solver_starts_as_created_and_matches_the_oracle():
    require_device()
    sp = configs.grin_fiber(M=6, nt=16384, symmetric=True)
    note("case autotune single M=6 NT=16384")
    plain = MultimodePropagationGPU_RamanON(sp, pair_decode=1)
    psi_plain, _ = plain.get_state()
    plain.close()
    a = MultimodePropagationGPU_RamanON(sp, pair_decode=1, flags=_capi.MMP_FLAG_AUTOTUNE)
    info = a.info()
    assert info["tuned"] == 1 and info["gather_waves"] in (4, 8) and info["chunk_lines"] in (16, 24, 40)
    assert info["steps_taken"] == 0 and info["z"] == 0.0
    psi, z = a.get_state()
    assert z == 0.0 and np.array_equal(psi, psi_plain)
    a.computeLawsonRK(20)
    ra = a.getResult()
    a.close()
    # the second solver of this structure takes the cached choice: the same arithmetic, bit for bit
    b = MultimodePropagationGPU_RamanON(sp, pair_decode=1, flags=_capi.MMP_FLAG_AUTOTUNE)
    ib = b.info()
    keys = ("gather_waves", "chunk_lines", "gather_hi_occupancy", "gather_matrix_cores", "pair_columns",
            "pair_groups")
    assert [ib[k] for k in keys] == [info[k] for k in keys]
    b.computeLawsonRK(20)
    assert np.array_equal(b.getResult(), ra)
    b.close()
    o = ob.Oracle(sp, pair_decode=1)
    o.compute_lawson_rk(20)
    assert np.all(configs.rel_l2_per_mode(ra, o.get_result()) <= 1e-12)


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
def test_tuned_batch_matches_untuned_and_takes_new_fibers(precision):
    """10 modes (several Raman-line chunks): the tuned batch against the untuned one, and a fiber
    loaded after tuning against the same fiber solved alone by an untuned solver."""
    require_device()
    fib = [configs.grin_fiber(M=10, nt=4096, n_steps=20000, fiber_length=0.4 + 0.1 * i, seed=30 + i,
                              symmetric=True, power_range=(1.0, 100.0)) for i in range(3)]
    note(f"case autotune batch M=10 NT=4096 precision={precision}")
    t = MultimodePropagationBatch(fib, pair_decode=1, precision=precision, flags=_capi.MMP_FLAG_AUTOTUNE)
    u = MultimodePropagationBatch(fib, pair_decode=1, precision=precision)
    assert t.info(0)["tuned"] == 1 and u.info(0)["tuned"] == 0
    new = configs.grin_fiber(M=10, nt=4096, n_steps=20000, fiber_length=0.3, seed=77, symmetric=True,
                             power_range=(1.0, 100.0))
    t.load(1, new)
    u.load(1, new)
    t.computeLawsonRK(16)
    u.computeLawsonRK(16)
    tol = 1e-12 if precision == _capi.MMP_PRECISION_FP64 else 1e-6
    for f in range(3):
        err = configs.rel_l2_per_mode(t.result(f), u.result(f))
        assert np.all(err <= tol), (f, err)
    t.close()
    u.close()


def test_tuning_leaves_no_fiber_parked():
    """A fiber whose field overflows during the tuning runs (1-10 MW pulses) is parked by the device
    there; the tuner restores its state and live flag, and it overflows again only when stepped."""
    require_device()
    sps = [configs.grin_fiber(M=3, nt=4096, n_steps=20000, seed=61, symmetric=True),
           configs.grin_fiber(M=3, nt=4096, n_steps=20000, fiber_length=1.0, seed=5, power_range=(1e6, 1e7),
                              symmetric=True)]
    note("case autotune with an overflowing fiber")
    b = MultimodePropagationBatch(sps, precision=_capi.MMP_PRECISION_MIXED, flags=_capi.MMP_FLAG_AUTOTUNE)
    for f in range(2):
        i = b.info(f)
        assert (i["halted"], i["live"], i["steps_taken"], i["z"]) == (0, 1, 0, 0.0), (f, i)
    b.computeLawsonRK(8)
    assert b.info(1)["halted"] == 1 and b.info(0)["halted"] == 0
    b.close()
