"""The column-stage gather with its coupling sums on the matrix cores (mixed precision): per
parity class of the coupling blocks, B[p][k] = sum over the class's slots of Qd[s][p][k] X_s as
FP32 v_mfma_f32_32x32x2_f32 products (rows (p, k), slots as the inner dimension, samples as the
columns), then sigma_p = sum_k U_k B[p][k] per sample. Against the FP64 oracle at the
mixed-precision bar of tests/test_gpu_mixed.py (1e-6, or the dynamics' amplification of an FP64
rounding-level perturbation scaled to FP32), and against the vector-ALU gather of the same
problem (MMP_FLAG_NO_MATRIX_CORES), which differs from it by FP32 rounding only. The default takes
the matrix cores at 4 to 10 and 13 to 16 modes whatever the number of fibers;
MMP_FLAG_MATRIX_CORES takes them at 11 and 12 modes too."""
import numpy as np
import pytest

import oracle_binding as ob
from gpu_util import note, require_device
from mm_propagation_amd import _capi, configs
from mm_propagation_amd.batch import MultimodePropagationBatch
from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON
from test_gpu_mixed import bar
from test_gpu_parity import break_parity

pytestmark = pytest.mark.gpu

MIXED = _capi.MMP_PRECISION_MIXED
VALU = _capi.MMP_FLAG_NO_MATRIX_CORES
MATRIX = _capi.MMP_FLAG_MATRIX_CORES


def solve(sp, steps, flags=0):
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=MIXED, flags=flags)
    info = s.info()
    s.computeLawsonRK(steps)
    r = s.getResult()
    s.close()
    return r, info


@pytest.mark.parametrize("M,nt,broken", [(4, 8192, True), (5, 4096, False), (6, 4096, True), (7, 4096, False),
                                         (8, 8192, False), (8, 4096, True), (9, 4096, False),
                                         (10, 16384, False), (10, 4096, True), (12, 4096, False), (15, 8192, False),
                                         (16, 4096, False)])
def test_matrix_core_gather_against_oracle_and_vector_gather(M, nt, broken):
    """5 to 16 modes with parity-compact coupling blocks (one to four 32-row tiles per class: up to
    two per wavefront with the wavefronts splitting the pairs, three or four with the wavefronts
    splitting the tiles; odd M with padding rows), and 4, 6, 8 and 10 modes with the parity
    structure broken (full blocks, one class, one to four tiles)."""
    require_device()
    sp = configs.grin_fiber(M=M, nt=nt, symmetric=True, power_range=(1.0, 100.0))
    if broken:
        sp = break_parity(sp)
    steps = 20
    note(f"case matrix-core gather M={M} NT={nt} broken={broken}")
    ra, ia = solve(sp, steps, MATRIX)
    rb, ib = solve(sp, steps, VALU)
    _, idef = solve(sp, 1)
    assert idef["gather_matrix_cores"] == (1 if M <= 10 or M >= 13 else 0), idef
    assert ia["gather_matrix_cores"] == 1 and ia["gather_waves"] == 4, ia
    assert ib["gather_matrix_cores"] == 0, ib
    o = ob.Oracle(sp, pair_decode=1, n_threads=16)
    o.compute_lawson_rk(steps)
    ref = o.get_result()
    o.close()
    b = bar(sp, steps, decode=1)
    ea, eb = configs.rel_l2_per_mode(ra, ref), configs.rel_l2_per_mode(rb, ref)
    print(f"M={M} matrix cores {ea.max():.2e}, vector ALUs {eb.max():.2e}, bar {b:.2e}", flush=True)
    assert np.all(ea <= b), ea
    assert np.all(eb <= b), eb
    assert np.all(configs.rel_l2_per_mode(ra, rb) <= 2 * b)


def test_matrix_core_batch_equals_single_solves():
    """Grid y = fiber: each fiber of a batch on the matrix-core gather is the same fiber solved
    alone, bit for bit."""
    require_device()
    sps = [configs.grin_fiber(M=10, nt=8192, n_steps=20000, fiber_length=0.5 + 0.1 * i, seed=120 + i,
                              symmetric=True) for i in range(3)]
    note("case matrix-core batch M=10 NT=8192")
    b = MultimodePropagationBatch(sps, pair_decode=1, precision=MIXED, flags=MATRIX)
    assert b.info(0)["gather_matrix_cores"] == 1
    b.computeLawsonRK(24)
    got = [b.result(f) for f in range(3)]
    b.close()
    for f, sp in enumerate(sps):
        r, _ = solve(sp, 24, MATRIX)
        assert np.array_equal(got[f], r), f


@pytest.mark.parametrize("M,nt,fibers,expected", [(10, 8192, 16, (0, 1)), (10, 8192, 1, (0, 1)), (6, 16384, 8, (0, 1)),
                                                  (6, 16384, 1, (0, 1)), (8, 16384, 8, (0, 1)), (4, 4096, 2, (0, 1)),
                                                  (12, 4096, 2, (0, 0)), (15, 4096, 2, (0, 1))])
def test_default_gather_variant(M, nt, fibers, expected):
    """(gather_hi_occupancy, gather_matrix_cores) chosen at creation: 4-10 and 13-16 modes take the
    matrix cores, 11 and 12 modes the vector gather, whatever the number of fibers."""
    require_device()
    sps = [configs.grin_fiber(M=M, nt=nt, seed=i, symmetric=True) for i in range(fibers)]
    b = MultimodePropagationBatch(sps, pair_decode=1, precision=MIXED)
    info = b.info(0)
    b.close()
    assert (info["gather_hi_occupancy"], info["gather_matrix_cores"]) == expected, info


@pytest.mark.parametrize("M,precision,broken", [(10, _capi.MMP_PRECISION_FP64, False), (3, MIXED, False),
                                                (11, MIXED, True), (20, MIXED, False)])
def test_no_matrix_cores_where_the_variant_does_not_exist(M, precision, broken):
    """FP64, fewer than 4 modes, more than four 32-row tiles per class (11 modes with full
    coupling blocks), or more than 16 modes (the separate passes): MMP_FLAG_MATRIX_CORES changes
    nothing there. With MMP_FLAG_NO_MATRIX_CORES as well it is refused."""
    require_device()
    sp = configs.grin_fiber(M=M, nt=4096 if M <= 16 else 256, symmetric=True)
    if broken:
        sp = break_parity(sp)
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=precision, flags=MATRIX)
    assert s.info()["gather_matrix_cores"] == 0
    s.close()
    with pytest.raises(_capi.MmpError):
        MultimodePropagationGPU_RamanON(sp, pair_decode=1, precision=precision, flags=MATRIX | VALU)
