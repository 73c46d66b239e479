"""The serving loop: requests of mixed shapes admitted into batch slots and solo solvers while
others run give, each of them, exactly the result of the same solve alone."""
import numpy as np
import pytest

from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.serving import Request, ServingEngine
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu


def mix():
    """8 requests: M = 3 and 6 at NT = 4096 (batch slots, 2 per group, so some wait), and NT = 2048
    (too short for the batched kernels: solo solvers); low powers so every field stays finite."""
    shapes = [(3, 4096, 40), (6, 4096, 56), (3, 4096, 24), (6, 2048, 30), (3, 4096, 64), (6, 4096, 20),
              (3, 2048, 45), (6, 4096, 33)]
    reqs = []
    for i, (M, nt, n) in enumerate(shapes):
        sp = configs.grin_fiber(M=M, nt=nt, n_steps=20000, fiber_length=0.3 + 0.05 * i, seed=50 + i,
                                t0=0.1 + 0.01 * i, symmetric=True)
        reqs.append(Request(rid=i, params=sp, n_steps=n, arrival=0.002 * i))
    return reqs


@pytest.mark.parametrize("precision", [_capi.MMP_PRECISION_MIXED, _capi.MMP_PRECISION_FP64])
def test_each_result_is_the_same_alone_and_in_the_mix(precision):
    require_device()
    reqs = mix()
    note(f"case serving mix precision={precision}")
    eng = ServingEngine(slots=2, max_chunk=16, precision=precision)
    stats = eng.run(reqs)
    eng.close()
    assert stats["requests"] == 8 and stats["nonfinite"] == 0
    assert any(r.placement.startswith("solo") for r in reqs) and any(r.placement.startswith("batch") for r in reqs)
    for r in reqs:
        s = MultimodePropagationGPU_RamanON(r.params, pair_decode=_capi.MMP_PAIRS_PHYSICAL, precision=precision)
        s.computeLawsonRK(r.n_steps)
        alone = s.getResult()
        s.close()
        assert np.array_equal(r.result, alone), (r.rid, r.placement)
        assert r.latency is not None and r.latency >= 0


def test_admission_waits_for_memory():
    """With a budget that fits one batch group at a time, the second group's requests wait in
    the queue and are admitted once memory frees; every request still completes."""
    require_device()
    reqs = [r for r in mix() if r.params.nt == 2048]  # two solo requests
    for r in reqs:
        r.arrival = 0.0
    from mm_propagation_amd import plan
    one = plan.plan(reqs[0].params, precision=_capi.MMP_PRECISION_MIXED, pair_decode=1)["device_bytes"]
    eng = ServingEngine(hbm_budget=1.5 * one, slots=2, max_chunk=8)
    stats = eng.run(reqs)
    eng.close()
    assert stats["requests"] == 2 and all(r.result is not None for r in reqs)
    assert stats["peak_device_bytes"] <= 1.5 * one


def test_idle_groups_are_reclaimed_for_other_structures():
    """A budget that holds one 2-slot batch group: requests of a second structure wait until
    the first group has no live request and is closed to make room; every result still equals
    its solve alone."""
    require_device()
    a = [Request(rid=i, params=configs.grin_fiber(M=3, nt=4096, n_steps=20000, fiber_length=0.3 + 0.1 * i,
                                                   seed=70 + i, symmetric=True), n_steps=24 + 8 * i)
         for i in range(2)]
    b = [Request(rid=2 + i, params=configs.grin_fiber(M=6, nt=4096, n_steps=20000, fiber_length=0.3 + 0.1 * i,
                                                       seed=80 + i, symmetric=True), n_steps=16 + 8 * i)
         for i in range(2)]
    from mm_propagation_amd import plan
    one_group = max(plan.plan(r.params, precision=_capi.MMP_PRECISION_FP64, pair_decode=1)["device_bytes"]
                    for r in a + b) * 2
    note("case serving reclaim")
    eng = ServingEngine(hbm_budget=1.2 * one_group, slots=2, max_chunk=8, precision=_capi.MMP_PRECISION_FP64)
    stats = eng.run(a + b)
    st = eng.stats()
    eng.close()
    assert stats["requests"] == 4 and stats["nonfinite"] == 0
    assert st["groups_created"] == 2 and st["groups"] == 1  # the first group was closed for the second
    assert stats["peak_device_bytes"] <= 1.2 * one_group
    for r in a + b:
        s = MultimodePropagationGPU_RamanON(r.params, pair_decode=_capi.MMP_PAIRS_PHYSICAL)
        s.computeLawsonRK(r.n_steps)
        assert np.array_equal(r.result, s.getResult()), r.rid
        s.close()


def test_autotuned_server_results_agree_with_solves_alone_to_rounding():
    """With MMP_FLAG_AUTOTUNE every group runs the gather variant tuned for its slot count, which
    may sum over the pair slots in another order than the solve alone: results agree to rounding."""
    require_device()
    reqs = [r for r in mix() if r.params.nt == 4096][:4]
    note("case serving with autotune, FP64")
    eng = ServingEngine(slots=2, max_chunk=16, precision=_capi.MMP_PRECISION_FP64, flags=_capi.MMP_FLAG_AUTOTUNE)
    stats = eng.run(reqs)
    eng.close()
    assert stats["requests"] == len(reqs) and stats["nonfinite"] == 0
    for r in reqs:
        s = MultimodePropagationGPU_RamanON(r.params, pair_decode=_capi.MMP_PAIRS_PHYSICAL)
        s.computeLawsonRK(r.n_steps)
        err = configs.rel_l2_per_mode(r.result, s.getResult())
        s.close()
        assert np.all(err <= 1e-12), (r.rid, err)


def test_overflowing_request_ends_early_and_frees_its_slot():
    """A request whose pulses (1-10 MW) overflow within a few steps is parked by the device and
    ends at the next chunk boundary as non-finite; the request that shares its group still
    gets the result of its solve alone."""
    require_device()
    ok = Request(rid=0, params=configs.grin_fiber(M=3, nt=4096, n_steps=20000, seed=71, symmetric=True), n_steps=40)
    bad = Request(rid=1, params=configs.grin_fiber(M=3, nt=4096, n_steps=20000, fiber_length=1.0, seed=5,
                                                   power_range=(1e6, 1e7), symmetric=True), n_steps=400)
    note("case serving overflowing request")
    eng = ServingEngine(slots=2, max_chunk=8, precision=_capi.MMP_PRECISION_FP64)
    stats = eng.run([ok, bad])
    eng.close()
    assert stats["nonfinite"] == 1 and bad.nonfinite and bad.result is None
    assert bad.finished is not None and bad.finished < ok.finished  # at its first chunk boundary, not after 400 steps
    s = MultimodePropagationGPU_RamanON(ok.params, pair_decode=_capi.MMP_PAIRS_PHYSICAL)
    s.computeLawsonRK(ok.n_steps)
    assert np.array_equal(ok.result, s.getResult())
    s.close()
