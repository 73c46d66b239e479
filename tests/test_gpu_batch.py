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


@pytest.mark.parametrize("M,flags,variant", [(7, 0, "matrix"), (7, _capi.MMP_FLAG_NO_MATRIX_CORES, "vector3"),
                                             (10, _capi.MMP_FLAG_NO_MATRIX_CORES, "vector3")])
def test_large_grid_variants_equal_single_solves(M, flags, variant):
    """16 fibers at NT = 8192 make 2048 gather workgroups: the batch runs a four-column pair pass
    with one line group, the Raman lines in column order before their row pass and the gather's
    large-grid variant (7 modes: the coupling sums on the matrix cores, or with
    MMP_FLAG_NO_MATRIX_CORES the three-workgroups-per-CU vector gather; 10 modes: the latter); a
    fiber alone four or eight line groups, the lines in row order and its small-grid gather
    (matrix cores at 7 modes, two-per-CU vector gather without them; at 10 modes the flag keeps
    the fiber alone off the matrix cores, which it would take on its small grid). The arithmetic of each
    gather is the same in the batch and alone, so the results are too."""
    require_device()
    F = 16
    sps = fibers(M, 8192, F, base=40)
    note(f"case large grid M={M} NT=8192 fibers={F} flags={flags}")
    b = MultimodePropagationBatch(sps, pair_decode=1, precision=_capi.MMP_PRECISION_MIXED, flags=flags)
    info = b.info(0)
    assert (info["pair_columns"], info["gather_waves"]) == (4, 4)
    assert (info["gather_hi_occupancy"], info["gather_matrix_cores"]) == ((0, 1) if variant == "matrix" else (1, 0))
    b.computeLawsonRK(24)
    got = {f: b.result(f) for f in (0, 9, F - 1)}
    b.close()
    for f, r in got.items():
        s = MultimodePropagationGPU_RamanON(sps[f], pair_decode=1, precision=_capi.MMP_PRECISION_MIXED, flags=flags)
        s.computeLawsonRK(24)
        assert np.array_equal(r, s.getResult()), f
        s.close()


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


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_FP64, _capi.MMP_PRECISION_MIXED])
def test_nonfinite_fiber_is_parked_by_the_device(precision):
    """A fiber whose pulses (1-10 MW) are far beyond what its step resolves overflows within a
    few steps: the step-closing kernel parks it (its z and step count stop, later steps skip it),
    the other fibers' results are those of their solves alone, resuming it is refused, and a new
    finite state makes it live again."""
    require_device()
    sps = fibers(3, 4096, 3, base=60)
    sps[1] = configs.grin_fiber(M=3, nt=4096, n_steps=20000, fiber_length=1.0, seed=5, power_range=(1e6, 1e7),
                                symmetric=True)
    note(f"case device parking precision={precision}")
    b = MultimodePropagationBatch(sps, precision=precision)
    psi1, _ = b.state(1)
    b.computeLawsonRK(32)
    i0, i1 = b.info(0), b.info(1)
    assert i1["halted"] == 1 and i1["live"] == 0 and 1 <= i1["steps_taken"] < 32, i1
    assert i0["halted"] == 0 and i0["steps_taken"] == 32
    assert b.state(1)[1] < b.state(0)[1]  # its z stopped
    with pytest.raises(_capi.MmpError) as ei:
        b.result(1)
    assert ei.value.status == _capi.MMP_ERR_NONFINITE
    with pytest.raises(_capi.MmpError) as ei:
        _capi.check(b.lib.mmp_fiber_set_live(b.h, 1, 1), "set_live")
    assert ei.value.status == _capi.MMP_ERR_NONFINITE
    got = [b.result(f) for f in (0, 2)]
    for f, r in zip((0, 2), got):
        assert np.array_equal(r, single(sps[f], 32, precision)), f
    # a finite state again: live, restarted at z = 0, and it steps with the others
    b.set_state(1, psi1 * 1e-3)
    i1 = b.info(1)
    assert i1["halted"] == 0 and i1["live"] == 1 and i1["steps_taken"] == 0
    b.advance(2)
    assert b.info(1)["steps_taken"] == 2
    b.close()


def test_single_solver_parks_a_nonfinite_field():
    """The same on a single solver: getResult reports MMP_ERR_NONFINITE and z stops at the step
    that overflowed (the reference would keep adding h to a NaN field)."""
    require_device()
    sp = configs.grin_fiber(M=3, nt=4096, n_steps=20000, fiber_length=1.0, seed=5, power_range=(1e6, 1e7),
                            symmetric=True)
    note("case single solver parking")
    s = MultimodePropagationGPU_RamanON(sp)
    s.computeLawsonRK(20)
    info = s.info()
    assert info["halted"] == 1 and info["steps_taken"] < 20
    with pytest.raises(_capi.MmpError) as ei:
        s.getResult()
    assert ei.value.status == _capi.MMP_ERR_NONFINITE
    s.close()


def test_group_of_mixed_shapes_equals_members_stepped_alone():
    """mmp_group: a mixed-precision batch (M = 3, NT = 4096), an FP64 batch (M = 6, NT = 8192) and a
    single FP64 fiber (M = 2, NT = 1024) on one context, stepped by one graph per 16 steps, give
    the results of each stepped on its own, bit for bit, and their z and step counts."""
    from mm_propagation_amd.batch import SolverGroup
    from mm_propagation_amd.solver import Context
    require_device()
    ctx = Context(0)
    a_sps = fibers(3, 4096, 3, base=90)
    b_sps = fibers(6, 8192, 2, base=95)
    c_sp = configs.two_mode_gaussian(nt=1024)
    note("case solver group of three shapes")

    def make():
        return (MultimodePropagationBatch(a_sps, precision=_capi.MMP_PRECISION_MIXED, ctx=ctx),
                MultimodePropagationBatch(b_sps, ctx=ctx),
                MultimodePropagationGPU_RamanON(c_sp, ctx=ctx))

    members = make()
    g = SolverGroup(members, ctx)
    g.advance(37)
    ctx.synchronize()
    got = [[members[0].result(f) for f in range(3)], [members[1].result(f) for f in range(2)], members[2].getResult()]
    zs = [members[0].info(0)["z"], members[1].info(1)["z"], members[2].info()["z"]]
    steps = [members[0].info(2)["steps_taken"], members[1].info(0)["steps_taken"], members[2].info()["steps_taken"]]
    g.close()
    alone = make()
    for m in alone:
        m.advance(37)
    ctx.synchronize()
    for f in range(3):
        assert np.array_equal(got[0][f], alone[0].result(f)), f
    for f in range(2):
        assert np.array_equal(got[1][f], alone[1].result(f)), f
    assert np.array_equal(got[2], alone[2].getResult())
    assert zs == [alone[0].info(0)["z"], alone[1].info(1)["z"], alone[2].info()["z"]]
    assert steps == [37, 37, 37]
    for m in members + alone:
        m.close()
    ctx.close()


def test_group_with_members_on_their_own_streams():
    """mmp_group whose members are on contexts (streams) of their own: each is a parallel branch of
    the group's graph. Group steps, a member's own steps in between and reads on the members'
    streams give the results of the same calls on the members alone, bit for bit."""
    from mm_propagation_amd.batch import SolverGroup
    from mm_propagation_amd.solver import Context
    require_device()
    gctx, ca, cb = Context(0), Context(0), Context(0)
    a_sps = fibers(3, 4096, 3, base=100)
    b_sps = fibers(6, 8192, 2, base=105)
    c_sp = configs.two_mode_gaussian(nt=1024)
    note("case solver group with parallel branches")

    def make():
        return (MultimodePropagationBatch(a_sps, precision=_capi.MMP_PRECISION_MIXED, ctx=ca),
                MultimodePropagationBatch(b_sps, ctx=cb),
                MultimodePropagationGPU_RamanON(c_sp, ctx=gctx))

    def results(ms):
        return [[ms[0].result(f) for f in range(3)], [ms[1].result(f) for f in range(2)], ms[2].getResult()]

    members = make()
    g = SolverGroup(members, gctx)
    g.advance(37)
    members[0].advance(3)  # on its own stream, after the group's steps
    g.advance(5)
    got = results(members)
    steps = [members[0].info(1)["steps_taken"], members[1].info(0)["steps_taken"], members[2].info()["steps_taken"]]
    g.close()
    alone = make()
    for m in alone:
        m.advance(37)
    alone[0].advance(3)
    for m in alone:
        m.advance(5)
    ref = results(alone)
    for f in range(3):
        assert np.array_equal(got[0][f], ref[0][f]), f
    for f in range(2):
        assert np.array_equal(got[1][f], ref[1][f]), f
    assert np.array_equal(got[2], ref[2])
    assert steps == [45, 42, 42]
    for m in members + alone:
        m.close()
    for c in (gctx, ca, cb):
        c.close()


def test_group_of_branches_only_with_parked_fibers():
    """Every member on a context of its own (the group's stream only forks and joins) and fibers
    that overflow on the way (parked by the device): group steps equal the members' own steps,
    bit for bit, parking included."""
    from mm_propagation_amd.batch import SolverGroup
    from mm_propagation_amd.solver import Context
    require_device()
    gctx, ca, cb = Context(0), Context(0), Context(0)
    a_sps = fibers(3, 4096, 3, base=130)
    a_sps[1] = configs.grin_fiber(M=3, nt=4096, n_steps=20000, fiber_length=1.0, seed=5, power_range=(1e6, 1e7),
                                  symmetric=True)  # overflows within a few steps
    b_sps = fibers(6, 4096, 2, base=135)
    note("case solver group of branches only, with a parked fiber")

    def make():
        return (MultimodePropagationBatch(a_sps, precision=_capi.MMP_PRECISION_MIXED, ctx=ca),
                MultimodePropagationBatch(b_sps, ctx=cb))

    members = make()
    g = SolverGroup(members, gctx)
    g.advance(0)
    g.advance(21)
    got = [[members[0].info(f)["halted"] for f in range(3)],
           [members[0].result(f) for f in (0, 2)], [members[1].result(f) for f in range(2)]]
    g.close()
    alone = make()
    for m in alone:
        m.advance(21)
    assert got[0] == [alone[0].info(f)["halted"] for f in range(3)] and got[0][1] == 1
    for r, f in zip(got[1], (0, 2)):
        assert np.array_equal(r, alone[0].result(f)), f
    for f in range(2):
        assert np.array_equal(got[2][f], alone[1].result(f)), f
    for m in members + alone:
        m.close()
    for c in (gctx, ca, cb):
        c.close()
