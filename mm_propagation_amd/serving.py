"""Continuous-batching serving of propagation solves: the Python face of the library's request
server (mmp_serve_* in include/mmprop.h; no reference counterpart).

Requests (a SimulationParameters and a step count, arriving over time) are submitted to the
native server when they arrive; the server admits them by device footprint into batch slots
(requests the batched kernels take, grouped by structure: n_modes, nt, tableau, coupling
pattern) or solvers of their own, keeps every group and solver on its own stream, refills
slots at chunk boundaries and parks empty ones. `ServingEngine.run` replays a list of
requests at their arrival times and reports p50/p99 latency and aggregate M*NT*steps/s.
A request's result equals the result of the same solve alone.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _capi, plan
from .problem import CParams, SimulationParameters


@dataclass
class Request:
    rid: int
    params: SimulationParameters
    n_steps: int
    arrival: float = 0.0  # seconds after the start of the run
    raman: bool = True
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[np.ndarray] = None
    nonfinite: bool = False
    placement: str = ""

    @property
    def units(self) -> int:
        return int(self.params.n_modes) * int(self.params.nt) * int(self.n_steps)

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


class ServingEngine:
    """One native request server (one device)."""

    def __init__(self, hbm_budget: float = 64e9, slots: int = 16, max_chunk: int = 256,
                 precision: int = _capi.MMP_PRECISION_MIXED, pair_decode: int = _capi.MMP_PAIRS_PHYSICAL,
                 flags: int = 0, poll_s: float = 2e-4, device: int = 0):
        self.lib = _capi.load_library()
        self.precision = precision
        self.pair_decode = pair_decode
        self.poll_s = poll_s
        opt = _capi.MmpServeOptions(device=device, slots=slots, max_chunk=max_chunk, precision=precision,
                                    pair_decode=pair_decode, flags=flags, hbm_budget=float(hbm_budget))
        h = ctypes.c_void_p()
        _capi.check(self.lib.mmp_serve_create(ctypes.byref(opt), ctypes.byref(h)), "mmp_serve_create")
        self.h = h
        self._plans: Dict[int, Dict[str, float]] = {}

    def plan_for(self, req: Request) -> Dict[str, float]:
        """mmp_plan of the request (host only): footprint and structure."""
        if req.rid not in self._plans:
            self._plans[req.rid] = plan.plan(req.params, raman=req.raman, pair_decode=self.pair_decode,
                                             precision=self.precision)
        return self._plans[req.rid]

    def submit(self, req: Request) -> int:
        cp = CParams(req.params, req.raman)
        rid = ctypes.c_uint64()
        _capi.check(self.lib.mmp_serve_submit(self.h, ctypes.byref(cp.c), int(req.n_steps), ctypes.byref(rid)),
                    "mmp_serve_submit")
        return rid.value

    def poll(self) -> int:
        active = ctypes.c_uint32()
        _capi.check(self.lib.mmp_serve_poll(self.h, ctypes.byref(active)), "mmp_serve_poll")
        return active.value

    def record(self, sid: int) -> _capi.MmpServeRecord:
        rec = _capi.MmpServeRecord()
        _capi.check(self.lib.mmp_serve_record_get(self.h, sid, ctypes.byref(rec)), "mmp_serve_record_get")
        return rec

    def result(self, sid: int, req: Request) -> Optional[np.ndarray]:
        out = np.empty((int(req.params.n_modes), int(req.params.nt)), dtype=np.complex128)
        st = self.lib.mmp_serve_result(self.h, sid, out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if st == _capi.MMP_ERR_NONFINITE:
            return None
        _capi.check(st, "mmp_serve_result")
        return out

    def stats(self) -> Dict[str, int]:
        st = _capi.MmpServeStats()
        _capi.check(self.lib.mmp_serve_stats_get(self.h, ctypes.byref(st)), "mmp_serve_stats_get")
        return {name: getattr(st, name) for name, _ in _capi.MmpServeStats._fields_}

    def run(self, requests: Sequence[Request], clock=time.perf_counter) -> Dict[str, float]:
        """Submits each request at its arrival time (seconds from now), polls the server until
        every request has finished, and fills in each request's result, times and placement."""
        pending = sorted(requests, key=lambda r: r.arrival)
        sids: List[int] = []
        submit_wall: List[float] = []
        t0 = clock()
        k = 0
        while True:
            now = clock() - t0
            while k < len(pending) and pending[k].arrival <= now:
                sids.append(self.submit(pending[k]))
                submit_wall.append(clock() - t0)
                k += 1
            active = self.poll()
            if k == len(pending) and active == 0:
                break
            if active == 0:
                time.sleep(max(0.0, pending[k].arrival - (clock() - t0)))
            else:
                time.sleep(self.poll_s)
        makespan = clock() - t0
        for req, sid, tw in zip(pending, sids, submit_wall):
            rec = self.record(sid)
            # server times are relative to its own clock; the request arrived at req.arrival
            req.admitted = tw + (rec.t_admit - rec.t_submit)
            req.finished = tw + (rec.t_finish - rec.t_submit)
            req.nonfinite = rec.state == _capi.MMP_REQ_NONFINITE
            req.result = None if req.nonfinite else self.result(sid, req)
            kind = f"batch group {rec.group} slot {rec.slot}" if rec.batched else "solo"
            req.placement = f"{kind} M={int(req.params.n_modes)} NT={int(req.params.nt)}"
        lat = np.array([r.latency for r in pending], dtype=np.float64)
        units = sum(r.units for r in pending)
        st = self.stats()
        return {"requests": len(pending), "makespan_s": makespan, "units_per_s": units / makespan,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "nonfinite": sum(1 for r in pending if r.nonfinite), "groups": st["groups_created"],
                "solos": st["solos_created"], "peak_device_bytes": st["peak_bytes"]}

    def close(self):
        if self.h:
            self.lib.mmp_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serving_requests(count: int, rate: float, seed: int = 0, modes=(1, 3, 6, 10, 15),
                     nts=(4096, 8192, 16384, 32768, 65536), steps=(500, 20000),
                     power_range=(10.0, 1.0e4)) -> List[Request]:
    """Poisson arrivals at `rate` requests/s (default_rng(seed)); per request: M and NT uniform
    over the given sets, n_steps log-uniform in `steps`, length U(0.2, 1) m, T0 U(0.05, 0.3) ps,
    peak powers log-uniform in `power_range`, GRIN coupling with seed = request id."""
    from . import configs
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        M = int(rng.choice(modes))
        nt = int(rng.choice(nts))
        n = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))))
        length = float(rng.uniform(0.2, 1.0))
        t0 = float(rng.uniform(0.05, 0.3))
        sp = configs.grin_fiber(M=M, nt=nt, n_steps=n, fiber_length=length, seed=1000 + i, t0=t0,
                                power_range=power_range, symmetric=True)
        out.append(Request(rid=i, params=sp, n_steps=n, arrival=t))
    return out
