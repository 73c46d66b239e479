"""Host side of the request server (no device needed): option checks, submission copies and
validates the request, records and results before any scheduling pass, the request generator."""
import ctypes

import numpy as np
import pytest

from mm_propagation_amd import _capi, configs
from mm_propagation_amd.problem import CParams
from mm_propagation_amd.serving import serving_requests


def server(lib, **kw):
    opt = dict(device=0, slots=4, max_chunk=64, precision=_capi.MMP_PRECISION_MIXED,
               pair_decode=_capi.MMP_PAIRS_PHYSICAL, flags=0, hbm_budget=1e9)
    opt.update(kw)
    h = ctypes.c_void_p()
    st = lib.mmp_serve_create(ctypes.byref(_capi.MmpServeOptions(**opt)), ctypes.byref(h))
    return st, h


def test_options_are_checked(capi_lib):
    for bad in (dict(slots=0), dict(max_chunk=0), dict(hbm_budget=0.0), dict(precision=7), dict(pair_decode=9)):
        st, h = server(capi_lib, **bad)
        assert st == _capi.MMP_ERR_INVALID_ARGUMENT and not h.value, bad


def test_submit_copies_and_validates(capi_lib):
    st, h = server(capi_lib)
    assert st == _capi.MMP_OK
    sp = configs.grin_fiber(M=3, nt=4096, n_steps=100, symmetric=True)
    cp = CParams(sp, True)
    rid = ctypes.c_uint64()
    assert capi_lib.mmp_serve_submit(h, ctypes.byref(cp.c), 100, ctypes.byref(rid)) == _capi.MMP_OK
    assert rid.value == 0
    del cp  # the server owns a copy
    assert capi_lib.mmp_serve_submit(h, ctypes.byref(CParams(sp, True).c), 0, None) == _capi.MMP_ERR_INVALID_ARGUMENT
    bad = configs.grin_fiber(M=3, nt=3000, n_steps=100, symmetric=True)  # nt not a power of two
    assert capi_lib.mmp_serve_submit(h, ctypes.byref(CParams(bad, True).c), 10, None) != _capi.MMP_OK
    rec = _capi.MmpServeRecord()
    assert capi_lib.mmp_serve_record_get(h, 0, ctypes.byref(rec)) == _capi.MMP_OK
    assert rec.state == _capi.MMP_REQ_QUEUED and rec.units == 3 * 4096 * 100 and rec.device_bytes > 0
    assert capi_lib.mmp_serve_record_get(h, 5, ctypes.byref(rec)) == _capi.MMP_ERR_INVALID_ARGUMENT
    out = np.empty(2 * 3 * 4096)
    assert capi_lib.mmp_serve_result(h, 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == _capi.MMP_ERR_STATE
    stats = _capi.MmpServeStats()
    assert capi_lib.mmp_serve_stats_get(h, ctypes.byref(stats)) == _capi.MMP_OK
    assert (stats.submitted, stats.queued, stats.running, stats.groups, stats.solos) == (1, 1, 0, 0, 0)
    capi_lib.mmp_serve_destroy(h)


def test_request_generator_is_seeded_and_in_range():
    a = serving_requests(20, 3.0, seed=4)
    b = serving_requests(20, 3.0, seed=4)
    assert [(r.params.n_modes, r.params.nt, r.n_steps, r.arrival) for r in a] == \
           [(r.params.n_modes, r.params.nt, r.n_steps, r.arrival) for r in b]
    assert all(r.params.n_modes in (1, 3, 6, 10, 15) and 4096 <= r.params.nt <= 65536 for r in a)
    assert all(500 <= r.n_steps <= 20000 for r in a)
    arr = [r.arrival for r in a]
    assert arr == sorted(arr) and arr[0] > 0


def test_request_beyond_the_budget_is_reported_not_waited_for(capi_lib):
    """A request whose footprint can never fit: drain returns MMP_ERR_STATE instead of waiting
    (admission refuses it before any device call)."""
    st, h = server(capi_lib, hbm_budget=1e5)
    sp = configs.grin_fiber(M=3, nt=4096, n_steps=100, symmetric=True)
    assert capi_lib.mmp_serve_submit(h, ctypes.byref(CParams(sp, True).c), 10, None) == _capi.MMP_OK
    assert capi_lib.mmp_serve_drain(h) == _capi.MMP_ERR_STATE
    rec = _capi.MmpServeRecord()
    capi_lib.mmp_serve_record_get(h, 0, ctypes.byref(rec))
    assert rec.state == _capi.MMP_REQ_QUEUED
    capi_lib.mmp_serve_destroy(h)
