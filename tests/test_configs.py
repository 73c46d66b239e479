"""Input generators: Blow-Wood response, grids, GRIN coupling rules."""
import math

import numpy as np

from mm_propagation_amd import configs


def test_blow_wood_closed_form():
    nt, tw = 1024, 10.0
    h = configs.blow_wood_response(nt, tw)
    t = np.arange(nt) * tw / nt
    t1, t2 = configs.TAU1, configs.TAU2
    expect = (t1 ** 2 + t2 ** 2) / (t1 * t2 ** 2) * np.exp(-t / t2) * np.sin(t / t1)
    assert np.allclose(h, expect, rtol=1e-13, atol=1e-13)
    assert h[0] == 0.0
    # normalised response: the integral of h_R over t is 1 (on a grid fine enough to resolve tau1)
    hf = configs.blow_wood_response(65536, tw)
    assert abs(hf.sum() * tw / 65536 - 1.0) < 1e-4


def test_time_grid_centre():
    t = configs.time_grid(8, 8.0)
    assert list(t) == [-4, -3, -2, -1, 0, 1, 2, 3]


def test_grin_rule_counts_and_parity():
    for M, nnz in ((6, 378), (10, 2750)):
        q = configs.grin_coupling(M, seed=0)
        assert sum(len(x) for x in q) == nnz
        for p in range(M):
            for nzv in q[p]:
                k, l, m = nzv.index
                assert k <= l and (k + l - m - p) % 2 == 0
                assert -0.5 * configs.SR <= nzv.value <= configs.SR
    assert sum(len(x) for x in configs.grin_coupling(6, 0, mirror=True)) == 648


def test_symmetric_rule_is_fully_symmetric():
    M = 5
    q = configs.grin_coupling_symmetric(M, np.random.default_rng(0))
    T = np.zeros((M,) * 4)
    for p in range(M):
        for nzv in q[p]:
            k, l, m = nzv.index
            T[p, k, l, m] = nzv.value
    for perm in ((1, 0, 2, 3), (0, 2, 1, 3), (3, 1, 2, 0), (2, 3, 0, 1)):
        assert np.array_equal(T, T.transpose(perm))
    assert sum(len(x) for x in q) == sum(1 for p in range(M) for k in range(M) for l in range(M) for m in range(M)
                                         if (k + l - m - p) % 2 == 0)


def test_named_workloads_shapes():
    sp = configs.two_mode_gaussian()
    assert (sp.n_modes, sp.nt, sp.n_steps, sp.fiber_length) == (2, 4096, 5000, 0.5)
    assert sp.initial_fields.shape == (2, 4096)
    mid = 2048
    assert abs(abs(sp.initial_fields[0, mid]) - math.sqrt(100)) < 1e-12
    g = configs.grin_fiber()
    assert (g.n_modes, g.nt, g.n_steps) == (6, 16384, 20000)
    assert len(g.dispersion_coefficients[0]) == 4
