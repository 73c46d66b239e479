"""Batched fibers (mmp_batch_create): every fiber of a batch equals the same fiber solved
alone, bit for bit (the batched kernels do the same arithmetic per fiber), and the oracle."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.batch import MultimodePropagationBatch
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu


def fibers(M, nt, count, base=0):
    """GRIN fibers of one structure with different coupling values, lengths, pulses."""
    out = []
    for i in range(count):
        sp = configs.grin_fiber(M=M, nt=nt, n_steps=20000, fiber_length=0.5 + 0.1 * i, seed=base + i,
                                t0=0.08 + 0.02 * i, symmetric=True)
        out.append(sp)
    return out


def single(sp, steps, precision=_capi.MMP_PRECISION_FP64):
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=precision)
    s.computeLawsonRK(steps)
    r = s.getResult()
    s.close()
    return r


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
@pytest.mark.parametrize("M,nt", [(3, 4096), (6, 8192), (10, 16384), (15, 4096)])
def test_batch_equals_single_solves(M, nt, precision):
    require_device()
    sps = fibers(M, nt, 3)
    note(f"case batch M={M} NT={nt} precision={precision}")
    b = MultimodePropagationBatch(sps, pair_decode=1, precision=precision)
    b.computeLawsonRK(40)
    got = [b.result(f) for f in range(3)]
    zs = [b.state(f)[1] for f in range(3)]
    b.close()
    for f, sp in enumerate(sps):
        s = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=precision)
        s.computeLawsonRK(40)
        assert np.array_equal(got[f], s.getResult()), f
        assert s.get_state()[1] == zs[f]
        s.close()


@pytest.mark.parametrize("M", [7, 10])
def test_large_grid_variants_equal_single_solves(M):
    """16 fibers at NT = 8192 make 2048 gather workgroups: the batch runs the three-workgroups-
    per-CU gather, a four-column pair pass with one line group and the Raman lines in column
    order before their row pass; a fiber alone the two-per-CU gather, four or eight line groups
    and the lines in row order. The arithmetic is the same, so the results are too."""
    require_device()
    F = 16
    sps = fibers(M, 8192, F, base=40)
    note(f"case large grid M={M} NT=8192 fibers={F}")
    b = MultimodePropagationBatch(sps, pair_decode=1, precision=_capi.MMP_PRECISION_MIXED)
    info = b.info(0)
    assert (info["pair_columns"], info["gather_waves"], info["gather_hi_occupancy"]) == (4, 4, 1)
    b.computeLawsonRK(24)
    got = {f: b.result(f) for f in (0, 9, F - 1)}
    b.close()
    for f, r in got.items():
        assert np.array_equal(r, single(sps[f], 24, _capi.MMP_PRECISION_MIXED)), f


def test_batch_against_oracle():
    require_device()
    sps = fibers(6, 4096, 2, base=7)
    note("case batch vs oracle")
    b = MultimodePropagationBatch(sps, pair_decode=1)
    b.computeLawsonRK(30)
    for f, sp in enumerate(sps):
        o = ob.Oracle(sp, pair_decode=1)
        o.compute_lawson_rk(30)
        assert np.all(configs.rel_l2_per_mode(b.result(f), o.get_result()) <= 1e-12)
        o.close()
    b.close()


def test_load_replaces_one_fiber_and_keeps_the_others():
    """A serving slot: fiber 1 is replaced after 16 steps by a new problem; it restarts at
    z = 0 while fibers 0 and 2 continue their solves."""
    require_device()
    sps = fibers(6, 4096, 3)
    new = configs.grin_fiber(M=6, nt=4096, n_steps=20000, fiber_length=0.3, seed=99, t0=0.15, symmetric=True)
    note("case batch load")
    b = MultimodePropagationBatch(sps, pair_decode=1)
    b.computeLawsonRK(16)
    b.load(1, new)
    b.advance(24)
    r = [b.result(f) for f in range(3)]
    steps = [b.info(f)["steps_taken"] for f in range(3)]
    b.close()
    assert steps == [40, 24, 40]
    assert np.array_equal(r[0], single(sps[0], 40))
    assert np.array_equal(r[2], single(sps[2], 40))
    s = MultimodePropagationGPU_RamanON(new, pair_decode=1)
    s.computeLawsonRK(16)  # its first steps, then 8 more: 24 in total, as in the batch
    s.advance(8)
    assert np.array_equal(r[1], s.getResult())
    s.close()


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
def test_parked_fiber_is_left_untouched_and_resumes(precision):
    """Fiber 1 is parked for 32 steps (graph replays and single steps): its state, z and step
    count stay as they were; resumed, it ends where 16 + 8 steps alone end. Fibers 0 and 2 are
    the 56-step single solves."""
    require_device()
    sps = fibers(3, 4096, 3, base=20)
    note(f"case batch park precision={precision}")
    b = MultimodePropagationBatch(sps, pair_decode=1, precision=precision)
    b.computeLawsonRK(16)
    psi0, z0 = b.state(1)
    b.set_live(1, False)
    b.advance(32)
    psi1, z1 = b.state(1)
    assert b.info(1)["live"] == 0 and b.info(1)["steps_taken"] == 16 and z1 == z0
    assert np.array_equal(psi0, psi1)
    b.set_live(1, True)
    b.advance(8)
    r = [b.result(f) for f in range(3)]
    steps = [b.info(f)["steps_taken"] for f in range(3)]
    b.close()
    assert steps == [56, 24, 56]
    assert np.array_equal(r[0], single(sps[0], 56, precision))
    assert np.array_equal(r[2], single(sps[2], 56, precision))
    assert np.array_equal(r[1], single(sps[1], 24, precision))


def test_batch_rejects_mixed_structures_and_ineligible_shapes():
    require_device()
    with pytest.raises(_capi.MmpError) as ei:
        MultimodePropagationBatch([configs.grin_fiber(M=3, nt=4096, symmetric=True),
                                   configs.grin_fiber(M=6, nt=4096, symmetric=True)], pair_decode=1)
    assert ei.value.status == _capi.MMP_ERR_INVALID_ARGUMENT
    with pytest.raises(_capi.MmpError) as ei:
        MultimodePropagationBatch([configs.grin_fiber(M=3, nt=2048, symmetric=True)] * 2, pair_decode=1)
    assert ei.value.status == _capi.MMP_ERR_UNSUPPORTED
