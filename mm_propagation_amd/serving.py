"""Continuous-batching serving loop for propagation solves (no reference counterpart).

Requests (a SimulationParameters and a step count, arriving over time) are admitted by device
memory footprint into

* batch slots: requests the batched kernels take (Raman on, 64-sample columns, M <= 16) are
  grouped by structure (n_modes, nt, tableau, coupling pattern) into one batched solver per
  group with a fixed number of slots; a new request is loaded into a free slot of an idle
  group (mmp_fiber_load) and joins the group's next chunk of steps;
* solo solvers: any other request gets a single-fiber solver of its own.

Every group and solo solver owns a context (one HIP stream), so their chunks overlap on the
device. A group advances by the smallest number of steps any of its active requests still
needs (capped), so no request takes a step too many; finished requests are read back and
their slots freed at chunk boundaries. Requests keep their fields and spectra on the device
for their whole solve. A request's result equals the result of the same solve alone: the
batched kernels do the same arithmetic per fiber and a solve split into chunks equals an
unsplit one.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi, plan
from .batch import MultimodePropagationBatch
from .problem import SimulationParameters
from .solver import Context, MultimodePropagationGPU_RamanOFF, MultimodePropagationGPU_RamanON


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


def _structure_key(req: Request, info: Dict[str, float]) -> Tuple:
    return (int(info["n_modes"]), int(info["nt"]), int(info["stages"]), int(info["n_pairs"]), int(info["nnz"]),
            int(info["n_conv_lines"]), req.raman)


class _Group:
    """A batched solver with fixed slots for requests of one structure."""

    def __init__(self, key, req: Request, slots: int, precision: int, pair_decode: int):
        self.key = key
        self.ctx = Context(0)
        self.batch = MultimodePropagationBatch([req.params] * slots, raman=req.raman, pair_decode=pair_decode,
                                               precision=precision, ctx=self.ctx)
        self.slots: List[Optional[Request]] = [None] * slots
        self.remaining = [0] * slots
        self.busy = False
        self.chunk = 0
        self.bytes = int(self.batch.info()["device_bytes"])
        for i in range(slots):
            self.batch.set_live(i, False)  # empty slots cost nothing until a request is loaded

    def free_slot(self) -> Optional[int]:
        for i, r in enumerate(self.slots):
            if r is None:
                return i
        return None

    def active(self) -> List[int]:
        return [i for i, r in enumerate(self.slots) if r is not None]

    def close(self):
        self.batch.close()
        self.ctx.close()


class _Solo:
    """A single-fiber solver for a request the batched kernels do not take."""

    def __init__(self, req: Request, precision: int, pair_decode: int):
        self.ctx = Context(0)
        cls = MultimodePropagationGPU_RamanON if req.raman else MultimodePropagationGPU_RamanOFF
        self.solver = cls(req.params, pair_decode=pair_decode, precision=precision, ctx=self.ctx)
        self.req = req
        self.remaining = req.n_steps
        self.started = False
        self.busy = False
        self.chunk = 0
        self.bytes = int(self.solver.info()["device_bytes"])

    def close(self):
        self.solver.close()
        self.ctx.close()


class ServingEngine:
    def __init__(self, hbm_budget: float = 64e9, slots: int = 16, max_chunk: int = 256,
                 precision: int = _capi.MMP_PRECISION_MIXED, pair_decode: int = _capi.MMP_PAIRS_PHYSICAL,
                 poll_s: float = 2e-4):
        self.hbm_budget = hbm_budget
        self.slots = slots
        self.max_chunk = max_chunk
        self.precision = precision
        self.pair_decode = pair_decode
        self.poll_s = poll_s
        self.groups: Dict[Tuple, _Group] = {}
        self.solos: List[_Solo] = []
        self.bytes_in_use = 0
        self.peak_bytes = 0
        self._plans: Dict[int, Dict[str, float]] = {}

    # --- placement --------------------------------------------------------
    def plan_for(self, req: Request) -> Dict[str, float]:
        if req.rid not in self._plans:  # host preparation once per request, not per admission attempt
            self._plans[req.rid] = plan.plan(req.params, raman=req.raman, pair_decode=self.pair_decode,
                                             precision=self.precision)
        return self._plans[req.rid]

    def _batchable(self, req: Request, info) -> bool:
        return req.raman and int(info["n1"]) == 64 and int(info["n_modes"]) <= 16

    def _admit(self, req: Request, now: float) -> bool:
        info = self.plan_for(req)
        if self._batchable(req, info):
            key = _structure_key(req, info)
            g = self.groups.get(key)
            if g is None:
                need = info["device_bytes"] * self.slots
                if self.bytes_in_use + need > self.hbm_budget:
                    return False
                g = _Group(key, req, self.slots, self.precision, self.pair_decode)
                self.groups[key] = g
                self.bytes_in_use += g.bytes
                self.peak_bytes = max(self.peak_bytes, self.bytes_in_use)
            if g.busy:
                return False  # slots are filled at chunk boundaries
            i = g.free_slot()
            if i is None:
                return False
            try:
                g.batch.load(i, req.params)
            except _capi.MmpError as e:
                if e.status != _capi.MMP_ERR_INVALID_ARGUMENT:
                    raise
                return self._admit_solo(req, info, now)  # same sizes, different coupling pattern
            g.slots[i] = req
            g.remaining[i] = req.n_steps
            req.admitted = now
            req.placement = f"batch M={key[0]} NT={key[1]} slot {i}"
            return True
        return self._admit_solo(req, info, now)

    def _admit_solo(self, req: Request, info, now: float) -> bool:
        if self.bytes_in_use + info["device_bytes"] > self.hbm_budget:
            return False
        s = _Solo(req, self.precision, self.pair_decode)
        self.solos.append(s)
        self.bytes_in_use += s.bytes
        self.peak_bytes = max(self.peak_bytes, self.bytes_in_use)
        req.admitted = now
        req.placement = f"solo M={int(info['n_modes'])} NT={int(info['nt'])}"
        return True

    # --- stepping ---------------------------------------------------------
    def _finish(self, req: Request, fields: Optional[np.ndarray], now: float):
        req.result = fields
        req.nonfinite = fields is None
        req.finished = now

    def _harvest_group(self, g: _Group, now: float):
        for i in g.active():
            g.remaining[i] -= g.chunk
            if g.remaining[i] == 0:
                req = g.slots[i]
                try:
                    fields = g.batch.result(i)
                except _capi.MmpError as e:
                    if e.status != _capi.MMP_ERR_NONFINITE:
                        raise
                    fields = None
                self._finish(req, fields, now)
                g.slots[i] = None
                g.batch.set_live(i, False)
        g.chunk = 0

    def _launch_group(self, g: _Group):
        act = g.active()
        if not act:
            return
        g.chunk = min(min(g.remaining[i] for i in act), self.max_chunk)
        g.batch.advance(g.chunk)
        g.busy = True

    def _step_solo(self, s: _Solo, now: float) -> bool:
        """Returns True when the solo request finished (the solver is then released)."""
        if s.busy:
            s.remaining -= s.chunk
            s.busy = False
        if s.remaining == 0:
            try:
                fields = s.solver.getResult()
            except _capi.MmpError as e:
                if e.status != _capi.MMP_ERR_NONFINITE:
                    raise
                fields = None
            self._finish(s.req, fields, now)
            return True
        s.chunk = min(s.remaining, self.max_chunk)
        if not s.started:
            s.solver.computeLawsonRK(s.chunk)  # the request's solve from z = 0
            s.started = True
        else:
            s.solver.advance(s.chunk)
        s.busy = True
        return False

    # --- the loop ---------------------------------------------------------
    def run(self, requests: Sequence[Request], clock=time.perf_counter) -> Dict[str, float]:
        pending = sorted(requests, key=lambda r: r.arrival)
        queue: List[Request] = []
        done = 0
        t0 = clock()
        k = 0
        while done < len(pending):
            now = clock() - t0
            while k < len(pending) and pending[k].arrival <= now:
                queue.append(pending[k])
                k += 1
            # chunk boundaries: harvest idle groups and solos, then admit, then launch
            for g in self.groups.values():
                if g.busy and g.ctx.idle():
                    g.busy = False
                    before = len(g.active())
                    self._harvest_group(g, clock() - t0)
                    done += before - len(g.active())
            for s in list(self.solos):
                if not s.busy or s.ctx.idle():
                    if self._step_solo(s, clock() - t0):
                        done += 1
                        self.solos.remove(s)
                        self.bytes_in_use -= s.bytes
                        s.close()
            for req in list(queue):
                if self._admit(req, clock() - t0):
                    queue.remove(req)
            for s in self.solos:
                if not s.busy and s.remaining > 0:
                    self._step_solo(s, clock() - t0)
            for g in self.groups.values():
                if not g.busy:
                    self._launch_group(g)
            busy = any(g.busy for g in self.groups.values()) or any(s.busy for s in self.solos)
            if not busy and not queue and k < len(pending):
                time.sleep(max(0.0, pending[k].arrival - (clock() - t0)))
            else:
                time.sleep(self.poll_s)
        makespan = clock() - t0
        lat = np.array([r.latency for r in pending], dtype=np.float64)
        units = sum(r.units for r in pending)
        return {"requests": len(pending), "makespan_s": makespan, "units_per_s": units / makespan,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "nonfinite": sum(1 for r in pending if r.nonfinite), "groups": len(self.groups),
                "peak_device_bytes": self.peak_bytes}

    def close(self):
        for g in self.groups.values():
            g.close()
        self.groups.clear()
        for s in self.solos:
            s.close()
        self.solos.clear()


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
