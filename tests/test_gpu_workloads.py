"""The benchmark workloads at their own shapes, against the FP64 oracle on identical inputs:
the 10-mode fiber at NT = 65536 (FP64 and mixed precision), the 256-fiber batched workload in
its nine (M, NT) class batches (mixed precision, the lowest-power fiber of every class, and FP64
batches of the lowest-power fibers of five classes), and a serving mix with 10- and 15-mode and
NT = 65536 requests.

Bars (per-mode rel-L2 against the oracle after the same steps):
  FP64_SHORT  = 1e-12: the plain FP64 path; short runs observe ~1e-15.
  MIXED_BAR   = 1e-6:  mixed precision, the FP32 target of the mixed-precision path.
  Mixed precision on a fiber whose dynamics amplify rounding quickly: the workload's pulses reach
  10 kW with as few as 10000 steps per metre, and on such a fiber an FP64 rounding-level
  perturbation (the oracle with its decimation-in-frequency FFT instead of decimation in time)
  grows by orders of magnitude within a few steps. There the bar is that growth scaled from FP64
  to FP32 rounding (2^29) times AMPLIFIED_FACTOR: a mixed-precision solver that rounds to FP32
  cannot stay closer than its own rounding amplified by the dynamics, and any error in the
  arithmetic (a wrong term, sign or index) shows as a difference of order one."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.batch import MultimodePropagationBatch
from mm_propagation_amd.serving import Request, ServingEngine
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON

pytestmark = pytest.mark.gpu

FP64_SHORT = 1e-12
MIXED_BAR = 1e-6
AMPLIFIED_FACTOR = 4.0
FP32_OVER_FP64 = 2.0 ** 29  # unit roundoff of FP32 over that of FP64
MIXED = _capi.MMP_PRECISION_MIXED
FP64 = _capi.MMP_PRECISION_FP64


def oracle_fields(sp, steps, fft_variant=0):
    o = ob.Oracle(sp, pair_decode=1, n_threads=16, fft_variant=fft_variant)
    o.compute_lawson_rk(steps)
    r = o.get_result()
    o.close()
    return r


def mixed_bar(sp, steps, ref):
    """MIXED_BAR, or the dynamics' amplification of an FP64 rounding-level perturbation over the
    same steps scaled to FP32 rounding when that is larger (module docstring)."""
    dif = oracle_fields(sp, steps, fft_variant=1)
    s64 = float(configs.rel_l2_per_mode(dif, ref).max())
    return max(MIXED_BAR, AMPLIFIED_FACTOR * FP32_OVER_FP64 * s64), s64


@pytest.mark.parametrize("precision,bar", [(FP64, FP64_SHORT), (MIXED, MIXED_BAR)])
def test_grin10_at_65536_samples(precision, bar):
    """The 10-mode GRIN fiber of the mixed-precision workload at its own grid (M = 10,
    NT = 65536, 10-1000 W per mode as written): 4 steps against the oracle."""
    require_device()
    sp = configs.grin_fiber(M=10, nt=65536, symmetric=True)
    steps = 4
    note(f"case grin10 NT=65536 precision={precision} steps={steps}")
    s = MultimodePropagationGPU_RamanON(sp, precision=precision)
    s.computeLawsonRK(steps)
    got = s.getResult()
    s.close()
    err = configs.rel_l2_per_mode(got, oracle_fields(sp, steps))
    print("rel-L2", err, flush=True)
    assert np.all(err <= bar), err


def batched_classes(count=256):
    fibers = [configs.batched_fiber(i) for i in range(count)]
    classes = {}
    for i, sp in enumerate(fibers):
        classes.setdefault((sp.n_modes, sp.nt), []).append(i)
    return fibers, classes


def peak_power(sp):
    return float(np.max(np.abs(sp.initial_fields) ** 2))


def test_batched_workload_every_class_against_oracle():
    """All 256 fibers of the batched workload, in their nine class batches (mixed precision),
    16 steps; in every class the lowest-power fiber is compared with the oracle."""
    require_device()
    fibers, classes = batched_classes()
    assert len(classes) == 9 and sum(len(v) for v in classes.values()) == 256
    steps = 16
    for key in sorted(classes):
        idx = classes[key]
        note(f"case batched class M={key[0]} NT={key[1]} fibers={len(idx)} mixed, {steps} steps")
        b = MultimodePropagationBatch([fibers[i] for i in idx], precision=MIXED)
        b.computeLawsonRK(steps)
        f = min(range(len(idx)), key=lambda j: peak_power(fibers[idx[j]]))
        got = b.result(f)
        b.close()
        sp = fibers[idx[f]]
        ref = oracle_fields(sp, steps)
        err = configs.rel_l2_per_mode(got, ref)
        bar, s64 = mixed_bar(sp, steps, ref)
        print(f"M={key[0]} NT={key[1]} fiber {idx[f]}: rel-L2 max {err.max():.3e} "
              f"(FP64 DIT/DIF {s64:.2e}, bar {bar:.2e})", flush=True)
        assert np.all(err <= bar), (key, idx[f], err, bar)


@pytest.mark.parametrize("key", [(3, 4096), (6, 8192), (10, 16384), (3, 16384), (10, 4096)])
def test_batched_workload_fp64_classes_against_oracle(key):
    """FP64 batches of workload fibers of one class (the four of lowest peak power: the most
    powerful pulses of the workload are not resolved by its step and leave the finite range
    within a few steps, in the oracle too), 8 steps; every fiber against the oracle at the FP64
    bar."""
    require_device()
    fibers, classes = batched_classes()
    idx = sorted(classes[key], key=lambda i: peak_power(fibers[i]))[:4]
    steps = 8
    note(f"case batched FP64 M={key[0]} NT={key[1]} fibers={idx}")
    b = MultimodePropagationBatch([fibers[i] for i in idx], precision=FP64)
    b.computeLawsonRK(steps)
    got = [b.result(f) for f in range(len(idx))]
    b.close()
    for f, i in enumerate(idx):
        err = configs.rel_l2_per_mode(got[f], oracle_fields(fibers[i], steps))
        assert np.all(err <= FP64_SHORT), (i, err)


def test_serving_mix_with_large_shapes_against_oracle():
    """A serving mix that holds 10- and 15-mode requests and NT = 65536 requests beside small
    ones (mixed precision, short step counts). Every result equals the same solve alone; the
    smallest requests, and the smallest of each mode count, are compared with the oracle."""
    require_device()
    shapes = [(15, 4096, 12), (10, 65536, 6), (3, 65536, 10), (1, 8192, 24), (6, 4096, 20), (10, 8192, 9),
              (15, 4096, 7), (3, 4096, 16), (1, 65536, 5), (6, 16384, 11)]
    reqs = []
    for i, (M, nt, n) in enumerate(shapes):
        sp = configs.grin_fiber(M=M, nt=nt, n_steps=20000, fiber_length=0.3 + 0.05 * i, seed=300 + i,
                                t0=0.08 + 0.02 * (i % 5), symmetric=True)
        reqs.append(Request(rid=i, params=sp, n_steps=n, arrival=0.003 * i))
    note("case serving mix with M = 10, 15 and NT = 65536 requests, mixed")
    eng = ServingEngine(slots=2, max_chunk=8, precision=MIXED)
    stats = eng.run(reqs)
    eng.close()
    assert stats["requests"] == len(reqs) and stats["nonfinite"] == 0
    for r in reqs:
        s = MultimodePropagationGPU_RamanON(r.params, precision=MIXED)
        s.computeLawsonRK(r.n_steps)
        assert np.array_equal(r.result, s.getResult()), (r.rid, r.placement)
        s.close()
    by_units = sorted(reqs, key=lambda r: r.units)
    checked = {r.rid: r for r in by_units[:3]}
    for M in (10, 15):
        r = min((r for r in reqs if r.params.n_modes == M and r.params.nt <= 8192), key=lambda r: r.units)
        checked[r.rid] = r
    for r in checked.values():
        note(f"case serving request {r.rid} M={r.params.n_modes} NT={r.params.nt} steps={r.n_steps} vs oracle")
        err = configs.rel_l2_per_mode(r.result, oracle_fields(r.params, r.n_steps))
        print(f"request {r.rid} ({r.placement}, {r.n_steps} steps): rel-L2 max {err.max():.3e}", flush=True)
        assert np.all(err <= MIXED_BAR), (r.rid, err)
