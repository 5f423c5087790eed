"""Serving many LES domains on one GPU.

Requests (a domain and a number of time steps) arrive over time; the server
admits a request when the device memory it needs fits in the budget, keeps
every admitted domain resident, and advances all of them concurrently: each
domain has its own stream and replays its captured step graph there, so the
steps of different domains overlap on the GPU.  A request's time to result
runs from its arrival to the completion of its last step on the device.

This is the host side only; each step is the library's whole-step path
(les_domain_step), unchanged.
"""
from __future__ import annotations

import dataclasses
import math
import time
from collections import deque

import numpy as np

from . import setup_gen


@dataclasses.dataclass
class Request:
    rid: int
    workload: setup_gen.Workload
    nsteps: int
    arrival: float = 0.0        # seconds after the start of the run

    @property
    def cells(self) -> int:
        w = self.workload
        return w.im * w.jm * w.km


@dataclasses.dataclass
class Result:
    rid: int
    cells: int
    nsteps: int
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def sample_requests(n: int, nsteps: int, rate: float | None = None, seed: int = 0,
                    sizes=None) -> list[Request]:
    """Mixed-size urban domains: nx, ny from {96, 128, ..., 512}, nz from
    {64, 96, 128}, building cover 20-45 %, median height 10-30 m
    (lognormal, sigma 0.5), u* 0.2-0.8 m/s, z0 0.05-0.5 m, dt 0.05 s.
    Arrivals form a Poisson process of the given rate (all at 0 when rate is
    None)."""
    rng = np.random.default_rng(seed)
    xs = np.arange(96, 513, 32)
    zs = np.array([64, 96, 128])
    out, t = [], 0.0
    for r in range(n):
        if sizes is not None:
            im, jm, km = sizes[r % len(sizes)]
        else:
            im, jm, km = int(rng.choice(xs)), int(rng.choice(xs)), int(rng.choice(zs))
        cover = rng.uniform(0.20, 0.45)
        city = setup_gen.CityParams(block=8, street=2, occupancy=min(1.0, cover / 0.64),
                                    median_h=float(rng.uniform(10.0, 30.0)), sigma=0.5)
        w = setup_gen.Workload(f"req{r}-{im}x{jm}x{km}", im, jm, km, city=city,
                               ustar=float(rng.uniform(0.2, 0.8)), z0=float(rng.uniform(0.05, 0.5)),
                               dt=0.05, nsteps=nsteps, seed=int(rng.integers(1 << 30)))
        if rate is not None and r > 0:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(r, w, nsteps, t if rate is not None else 0.0))
    return out


class _Active:
    def __init__(self, req: Request, dom, admitted: float):
        self.req = req
        self.dom = dom
        self.admitted = admitted
        self.issued = 0


class Server:
    """Admission by device memory, concurrent stepping, completion by polling.

    domain_factory(request) -> object with step(n, nmax), done(), close();
    bytes_of(request) -> device bytes the request's domain holds.  The
    defaults create library domains from the light set-up (masks and cn1
    derived on the device).
    """

    def __init__(self, budget_bytes: int, burst: int = 4, max_active: int = 64, domain_factory=None,
                 bytes_of=None, clock=time.perf_counter, batch_factory=None):
        self.budget = int(budget_bytes)
        self.burst = int(burst)
        self.max_active = int(max_active)
        self.clock = clock
        if domain_factory is None or bytes_of is None:
            from . import api

            def factory(req):
                d = api.Domain(setup_gen.build(req.workload, dense=False))
                d.set_graph(True)
                d.autotune()  # SOR launch shape for this grid shape (timed once per shape)
                return d

            domain_factory = domain_factory or factory
            bytes_of = bytes_of or (lambda req: api.bytes_needed(req.workload.im, req.workload.jm,
                                                                req.workload.km))
        self.factory = domain_factory
        self.bytes_of = bytes_of
        if batch_factory is None:
            def batch_factory(doms):
                from . import api
                return api.Batch(doms)
        # batch_factory(domains) -> object with step(ns, nmaxs), done(), close()
        self.batch_factory = batch_factory

    def run_batched(self, requests) -> list[Result]:
        """Continuous batching: the admitted domains advance in lockstep,
        one batched step (one launch per stage and per SOR iteration for all
        of them) at a time; requests join at step boundaries as memory frees
        up, and leave when their last step has completed."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        for r in pending:
            if self.bytes_of(r) > self.budget:
                raise ValueError(f"request {r.rid} needs more device memory than the whole budget")
        active: list[_Active] = []
        batch = None
        used = 0
        results = []
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            joined = False
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                need = self.bytes_of(pending[0])
                if used + need > self.budget:
                    break
                req = pending.popleft()
                active.append(_Active(req, self.factory(req), self.clock() - t0))
                used += need
                joined = True
            if not active:
                wait = pending[0].arrival - (self.clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
                continue
            if joined or batch is None:
                if batch is not None:
                    batch.close()
                batch = self.batch_factory([a.dom for a in active])
            k = min(self.burst, min(a.req.nsteps - a.issued for a in active))
            for _ in range(k):
                batch.step([a.issued + 1 for a in active], [a.req.nsteps for a in active])
                for a in active:
                    a.issued += 1
            if any(a.issued >= a.req.nsteps for a in active):
                while not batch.done():
                    time.sleep(20e-6)
                still = []
                for a in active:
                    if a.issued >= a.req.nsteps:
                        results.append(Result(a.req.rid, a.req.cells, a.req.nsteps, a.req.arrival, a.admitted,
                                              self.clock() - t0))
                        used -= self.bytes_of(a.req)
                    else:
                        still.append(a)
                batch.close()
                batch = None
                for a in active:
                    if a.issued >= a.req.nsteps:
                        a.dom.close()
                active = still
        if batch is not None:
            batch.close()
        return sorted(results, key=lambda r: r.rid)

    def run(self, requests, sequential: bool = False) -> list[Result]:
        """Serve the requests (arrival times relative to now); with
        sequential=True one domain at a time, each run to completion."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        for r in pending:
            if self.bytes_of(r) > self.budget:
                raise ValueError(f"request {r.rid} needs more device memory than the whole budget")
        active: list[_Active] = []
        used = 0
        results = []
        t0 = self.clock()
        cap = 1 if sequential else self.max_active
        while pending or active:
            now = self.clock() - t0
            # admission: first come, first served, while the memory fits
            while pending and pending[0].arrival <= now and len(active) < cap:
                need = self.bytes_of(pending[0])
                if used + need > self.budget:
                    break
                req = pending.popleft()
                active.append(_Active(req, self.factory(req), self.clock() - t0))
                used += need
            progressed = False
            for a in active:
                if a.issued < a.req.nsteps:
                    n1 = min(a.req.nsteps, a.issued + (a.req.nsteps if sequential else self.burst))
                    for n in range(a.issued + 1, n1 + 1):
                        a.dom.step(n, a.req.nsteps)
                    a.issued = n1
                    progressed = True
            still = []
            for a in active:
                if a.issued >= a.req.nsteps and a.dom.done():
                    results.append(Result(a.req.rid, a.req.cells, a.req.nsteps, a.req.arrival, a.admitted,
                                          self.clock() - t0))
                    a.dom.close()
                    used -= self.bytes_of(a.req)
                else:
                    still.append(a)
            active = still
            if not progressed and active:
                time.sleep(20e-6)
            elif not active and pending:
                wait = pending[0].arrival - (self.clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        return sorted(results, key=lambda r: r.rid)


def summarize(results: list[Result]) -> dict:
    """Aggregate cell updates per second over the run and the latency
    percentiles."""
    if not results:
        return {}
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    updates = sum(r.cells * r.nsteps for r in results)
    lat = np.array([r.latency for r in results])
    return {
        "requests": len(results),
        "mcell_updates_per_s": updates / span / 1e6 if span > 0 else math.inf,
        "span_s": span,
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
    }
