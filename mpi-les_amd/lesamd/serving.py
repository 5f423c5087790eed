"""Serving many LES domains on one GPU.

Requests (a domain and a number of time steps) arrive over time; the server
admits a request when the device memory it needs fits in the budget, keeps
every admitted domain resident, and advances all of them concurrently:
either each domain on its own stream, replaying its captured step graph
there (run), or the smaller domains in lockstep as one batch, every stage one
launch for all of them, with the larger ones on their own streams alongside
(run_batched).  Domains are built and released on a worker thread, so that
the steps of the others keep being issued meanwhile.  A request's time to
result runs from its arrival to the completion of its last step on the
device.

This is the host side only; each step is the library's whole-step or
batched-step path (les_domain_step, les_batch_step), unchanged.
"""
from __future__ import annotations

import dataclasses
import math
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

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
    # "ok", or "failed: ..." when the domain's state was not finite at its
    # last pressure solve (the simulation diverged).  A failed request fails
    # alone: the others sharing the device and the batch are unaffected.
    status: str = "ok"

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def ok(self) -> bool:
        return self.status == "ok"


def _status(dom) -> str:
    """The request's outcome from its finished domain: a non-finite
    pressure-solve residual (the SOR's stop-test sum, press.f95:119-121, is
    NaN or infinite) marks a diverged state.  Domain objects without the
    check (stand-ins) count as ok."""
    check = getattr(dom, "nonfinite", None)
    if check is not None and check():
        return "failed: non-finite state (diverged)"
    return "ok"


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


class _Lifecycle:
    """Admission and the domains' life cycle for one run: first come, first
    served, a request is admitted once it has arrived and the device memory
    its domain holds fits the budget.  With a worker thread the domain is
    built there and, when its request is done, closed there: creating and
    releasing device memory and uploading the set-up wait for the device, and
    must not hold up the issue of the other domains' steps.  The memory stays
    counted from admission until the close has returned."""

    def __init__(self, srv: "Server", requests, cap: int, t0: float):
        self.srv, self.cap, self.t0 = srv, cap, t0
        self.pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        for r in self.pending:
            if srv.bytes_of(r) > srv.budget:
                raise ValueError(f"request {r.rid} needs more device memory than the whole budget")
        self.pool = ThreadPoolExecutor(max_workers=srv.workers) if srv.workers > 0 else None
        self.creating: deque = deque()
        self.closing: list = []
        self.used = 0

    def now(self) -> float:
        return self.srv.clock() - self.t0

    def admit(self, nactive: int) -> list[_Active]:
        """Start the domains of the requests now admissible; return those
        whose domain is ready (in admission order)."""
        ready = []
        now = self.now()
        while (self.pending and self.pending[0].arrival <= now
               and nactive + len(self.creating) + len(ready) < self.cap):
            need = self.srv.bytes_of(self.pending[0])
            if self.used + need > self.srv.budget:
                break
            req = self.pending.popleft()
            self.used += need
            if self.pool is None:
                ready.append(_Active(req, self.srv.factory(req), self.now()))
            else:
                self.creating.append((req, self.pool.submit(self.srv.factory, req)))
        while self.creating and self.creating[0][1].done():
            req, fut = self.creating.popleft()
            ready.append(_Active(req, fut.result(), self.now()))
        return ready

    def release(self, a: _Active) -> None:
        need = self.srv.bytes_of(a.req)
        if self.pool is None:
            a.dom.close()
            self.used -= need
        else:
            self.closing.append((self.pool.submit(a.dom.close), need))

    def reap(self) -> bool:
        """Account the closes that have returned; True if any had."""
        keep, reaped = [], False
        for fut, need in self.closing:
            if fut.done():
                fut.result()
                self.used -= need
                reaped = True
            else:
                keep.append((fut, need))
        self.closing = keep
        return reaped

    def waiting(self) -> bool:
        return bool(self.pending or self.creating)

    def idle_wait(self, busy: bool) -> None:
        """Nothing was issued: poll again shortly, or sleep until the next
        arrival when nothing is in flight."""
        if busy or self.creating or self.closing:
            time.sleep(20e-6)
        elif self.pending:
            wait = self.pending[0].arrival - self.now()
            if wait > 0:
                time.sleep(min(wait, 1e-3))

    def finish(self) -> None:
        if self.pool is not None:
            self.pool.shutdown(wait=True)
            for fut, need in self.closing:
                fut.result()
                self.used -= need
            self.closing = []


class Server:
    """Admission by device memory, concurrent stepping, completion by polling.

    domain_factory(request) -> object with step(n, nmax), done(), close();
    bytes_of(request) -> device bytes the request's domain holds.  The
    defaults create library domains from the light set-up (masks and cn1
    derived on the device).  workers > 0 builds and closes the domains on
    that many worker threads (0: on the serving thread itself).
    """

    def __init__(self, budget_bytes: int, burst: int = 4, max_active: int = 64, domain_factory=None,
                 bytes_of=None, clock=time.perf_counter, batch_factory=None, workers: int = 1):
        self.budget = int(budget_bytes)
        self.burst = int(burst)
        self.max_active = int(max_active)
        self.clock = clock
        self.workers = int(workers)
        self.on_result = None  # called as on_result(done, total) after each completed request
        if domain_factory is None or bytes_of is None:
            from . import api

            def factory(req):
                d = api.Domain(setup_gen.build(req.workload, dense=False))
                d.set_graph(True)
                d.autotune()  # SOR launch shape and stencil k levels for this grid shape (timed once per shape)
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

    def warm(self, requests, progress=None) -> None:
        """Create (and close) one domain of every grid shape among the
        requests, untimed: the default factory's per-shape tuning (SOR launch
        shape, stencil k levels) is timed once per shape and process, so
        after this every serving mode starts from the same tuned state.
        progress(done, total) is called after each shape."""
        shapes = []
        for r in requests:
            w = r.workload
            if (w.im, w.jm, w.km) not in {(x.workload.im, x.workload.jm, x.workload.km) for x in shapes}:
                shapes.append(r)
        for q, r in enumerate(shapes):
            self.factory(r).close()
            if progress is not None:
                progress(q + 1, len(shapes))

    def run_batched(self, requests, max_batch_cells: int | None = None) -> list[Result]:
        """Continuous batching: the admitted domains of at most
        max_batch_cells cells (all of them when None) advance in lockstep, one
        batched step (one launch per stage and per SOR iteration for all of
        them) at a time, on the batch's stream; larger domains, which fill the
        GPU by themselves, replay their own step graphs on their own streams
        alongside.  Requests join at step boundaries once their domain is
        ready, and leave when their last step has completed."""
        lc = _Lifecycle(self, requests, self.max_active, self.clock())
        batched: list[_Active] = []
        alone: list[_Active] = []
        joining: list[_Active] = []
        batch = None
        results = []

        def done_with(a):
            results.append(Result(a.req.rid, a.req.cells, a.req.nsteps, a.req.arrival, a.admitted, lc.now(),
                                  _status(a.dom)))
            lc.release(a)
            if self.on_result is not None:
                self.on_result(len(results), len(requests))

        try:
            while lc.waiting() or batched or alone or joining:
                progressed = lc.reap()
                for a in lc.admit(len(batched) + len(alone) + len(joining)):
                    if max_batch_cells is None or a.req.cells <= max_batch_cells:
                        joining.append(a)
                    else:
                        alone.append(a)
                # the batch: members leave once the batch's stream has drained,
                # new members join between batched steps
                if batched and any(a.issued >= a.req.nsteps for a in batched):
                    if batch is None or batch.done():
                        if batch is not None:
                            batch.close()
                            batch = None
                        for a in batched:
                            if a.issued >= a.req.nsteps:
                                done_with(a)
                        batched = [a for a in batched if a.issued < a.req.nsteps]
                        progressed = True
                elif batched or joining:
                    if joining or batch is None:
                        if batch is not None:
                            batch.close()
                        batched += joining
                        joining = []
                        batch = self.batch_factory([a.dom for a in batched])
                    k = min(self.burst, min(a.req.nsteps - a.issued for a in batched))
                    for _ in range(k):
                        batch.step([a.issued + 1 for a in batched], [a.req.nsteps for a in batched])
                        for a in batched:
                            a.issued += 1
                    progressed = True
                # the domains on their own streams
                still = []
                for a in alone:
                    if a.issued < a.req.nsteps:
                        n1 = min(a.req.nsteps, a.issued + self.burst)
                        for n in range(a.issued + 1, n1 + 1):
                            a.dom.step(n, a.req.nsteps)
                        a.issued = n1
                        progressed = True
                    if a.issued >= a.req.nsteps and a.dom.done():
                        done_with(a)
                        progressed = True
                    else:
                        still.append(a)
                alone = still
                if not progressed:
                    lc.idle_wait(bool(batched or alone))
        finally:
            if batch is not None:
                batch.close()
            lc.finish()
        return sorted(results, key=lambda r: r.rid)

    def run(self, requests, sequential: bool = False) -> list[Result]:
        """Serve the requests (arrival times relative to now); with
        sequential=True one domain at a time, each run to completion."""
        lc = _Lifecycle(self, requests, 1 if sequential else self.max_active, self.clock())
        active: list[_Active] = []
        results = []
        try:
            while lc.waiting() or active:
                progressed = lc.reap()
                active += lc.admit(len(active))
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
                                              lc.now(), _status(a.dom)))
                        lc.release(a)
                        progressed = True
                        if self.on_result is not None:
                            self.on_result(len(results), len(requests))
                    else:
                        still.append(a)
                active = still
                if not progressed:
                    lc.idle_wait(bool(active))
        finally:
            lc.finish()
        return sorted(results, key=lambda r: r.rid)


def summarize(results: list[Result]) -> dict:
    """Aggregate cell updates per second over the run and the latency
    percentiles of the requests that completed; failed requests are counted
    apart and add no cell updates (a diverged domain's later solves exit at
    once, so its steps would inflate the rate)."""
    if not results:
        return {}
    good = [r for r in results if r.ok]
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    updates = sum(r.cells * r.nsteps for r in good)
    lat = np.array([r.latency for r in good]) if good else np.array([math.nan])
    return {
        "requests": len(results),
        "failed": len(results) - len(good),
        "mcell_updates_per_s": updates / span / 1e6 if span > 0 else math.inf,
        "span_s": span,
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
    }
