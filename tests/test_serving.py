"""The serving scheduler (admission by device memory, concurrent stepping,
completion by polling) on stand-in domains: no GPU needed."""
import itertools

import pytest

from lesamd import serving


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4
        return self.t


class FakeDomain:
    live = 0
    peak = 0

    def __init__(self, req, log):
        self.req, self.log, self.steps, self.polls = req, log, [], 0
        self.closed = False
        FakeDomain.live += 1
        FakeDomain.peak = max(FakeDomain.peak, FakeDomain.live)

    def step(self, n, nmax):
        # the life cycle never steps, polls or closes a closed domain
        assert not self.closed, f"step {n} issued on closed domain {self.req.rid}"
        assert nmax == self.req.nsteps
        self.steps.append(n)

    def done(self):
        assert not self.closed
        self.polls += 1
        return self.polls >= 2

    def close(self):
        assert not self.closed, f"domain {self.req.rid} closed twice"
        self.closed = True
        FakeDomain.live -= 1
        self.log.append((self.req.rid, tuple(self.steps)))


def make(budget, reqs, sizes, burst=3, max_active=64, workers=1):
    log = []
    srv = serving.Server(budget, burst=burst, max_active=max_active,
                         domain_factory=lambda r: FakeDomain(r, log), bytes_of=lambda r: sizes[r.rid],
                         clock=FakeClock(), workers=workers)
    return srv, log


def _reqs(n, nsteps=7):
    return [serving.Request(r, serving.setup_gen.Workload(f"r{r}", 8, 8, 4), nsteps) for r in range(n)]


@pytest.mark.parametrize("workers", [0, 1])
def test_every_request_runs_every_step_in_order(workers):
    FakeDomain.live = FakeDomain.peak = 0
    reqs = _reqs(5)
    srv, log = make(10 ** 9, reqs, {r.rid: 100 for r in reqs}, workers=workers)
    res = srv.run(reqs)
    assert [r.rid for r in res] == list(range(5))
    assert sorted(log) == [(r, tuple(range(1, 8))) for r in range(5)]
    assert all(r.finished >= r.admitted >= r.arrival for r in res)
    assert FakeDomain.peak == 5


@pytest.mark.parametrize("workers", [0, 1])
def test_admission_never_exceeds_the_memory_budget(workers):
    FakeDomain.live = FakeDomain.peak = 0
    reqs = _reqs(6)
    sizes = {0: 60, 1: 50, 2: 40, 3: 30, 4: 70, 5: 10}
    srv, log = make(100, reqs, sizes, workers=workers)
    srv.run(reqs)
    # first come first served: 0 (60) then 1 would make 110 > 100, so only 0 runs alone
    assert FakeDomain.peak <= 3
    assert len(log) == 6


def test_request_larger_than_budget_is_refused():
    reqs = _reqs(1)
    srv, _ = make(10, reqs, {0: 11})
    with pytest.raises(ValueError):
        srv.run(reqs)


@pytest.mark.parametrize("workers", [0, 1])
def test_sequential_mode_runs_one_domain_at_a_time(workers):
    FakeDomain.live = FakeDomain.peak = 0
    reqs = _reqs(4)
    srv, log = make(10 ** 9, reqs, {r.rid: 1 for r in reqs}, workers=workers)
    srv.run(reqs, sequential=True)
    assert FakeDomain.peak == 1
    assert [r for r, _ in log] == [0, 1, 2, 3]


def test_sampled_requests_follow_the_mixed_size_distribution():
    reqs = serving.sample_requests(200, 10, rate=5.0, seed=3)
    dims = {(r.workload.im, r.workload.jm, r.workload.km) for r in reqs}
    assert all(96 <= im <= 512 and im % 32 == 0 and km in (64, 96, 128) for im, _, km in dims)
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr) and arr[0] == 0.0
    gaps = [b - a for a, b in itertools.pairwise(arr)] if hasattr(itertools, "pairwise") else \
        [arr[i + 1] - arr[i] for i in range(len(arr) - 1)]
    assert 0.1 < sum(gaps) / len(gaps) < 0.3   # mean inter-arrival 1/rate
    again = serving.sample_requests(200, 10, rate=5.0, seed=3)
    assert [r.workload for r in again] == [r.workload for r in reqs]


def test_summary_percentiles():
    res = [serving.Result(i, 1000, 10, 0.0, 0.0, float(i + 1)) for i in range(100)]
    s = serving.summarize(res)
    assert s["requests"] == 100 and abs(s["p50_s"] - 50.5) < 1e-9 and s["p99_s"] > 99
    assert abs(s["mcell_updates_per_s"] - 100 * 1000 * 10 / 100 / 1e6) < 1e-12


class FakeBatch:
    def __init__(self, doms, blog):
        assert not any(d.closed for d in doms), "batch built from a closed domain"
        self.doms, self.blog, self.polls = doms, blog, 0
        self.closed = False
        blog.append(("open", tuple(d.req.rid for d in doms)))

    def step(self, ns, nmaxs):
        assert len(ns) == len(self.doms) == len(nmaxs)
        for d, n, m in zip(self.doms, ns, nmaxs):
            d.step(n, m)

    def done(self):
        self.polls += 1
        return self.polls >= 2

    def close(self):
        # a batch is closed before any of its domains (they must outlive it)
        assert not self.closed and not any(d.closed for d in self.doms), "batch outlived one of its domains"
        self.closed = True
        self.blog.append(("close", tuple(d.req.rid for d in self.doms)))


@pytest.mark.parametrize("workers", [0, 1])
def test_batched_serving_steps_every_request_in_order_within_the_budget(workers):
    FakeDomain.live = FakeDomain.peak = 0
    reqs = [serving.Request(r, serving.setup_gen.Workload(f"r{r}", 8, 8, 4), 5 + r) for r in range(6)]
    sizes = {0: 60, 1: 30, 2: 40, 3: 30, 4: 70, 5: 10}
    log, blog = [], []
    srv = serving.Server(100, burst=3, domain_factory=lambda r: FakeDomain(r, log),
                         bytes_of=lambda r: sizes[r.rid], clock=FakeClock(),
                         batch_factory=lambda doms: FakeBatch(doms, blog), workers=workers)
    res = srv.run_batched(reqs)
    assert [r.rid for r in res] == list(range(6))
    assert sorted(log) == [(r, tuple(range(1, 6 + r))) for r in range(6)]
    # every batch holds admitted requests within the budget, and each is closed
    opens = [b for kind, b in blog if kind == "open"]
    assert all(sum(sizes[r] for r in b) <= 100 for b in opens)
    assert len(opens) == sum(1 for kind, _ in blog if kind == "close")
    assert FakeDomain.live == 0


@pytest.mark.parametrize("workers", [0, 1])
def test_batched_serving_steps_large_domains_on_their_own(workers):
    FakeDomain.live = FakeDomain.peak = 0
    # requests 1, 3, 5 have 512 cells, the others 256
    reqs = [serving.Request(r, serving.setup_gen.Workload(f"r{r}", 8, 8, 4 * (1 + r % 2)), 4 + r) for r in range(6)]
    log, blog = [], []
    srv = serving.Server(1000, burst=2, domain_factory=lambda r: FakeDomain(r, log),
                         bytes_of=lambda r: 10, clock=FakeClock(),
                         batch_factory=lambda doms: FakeBatch(doms, blog), workers=workers)
    res = srv.run_batched(reqs, max_batch_cells=256)
    assert [r.rid for r in res] == list(range(6))
    assert sorted(log) == [(r, tuple(range(1, 5 + r))) for r in range(6)]
    opens = [b for kind, b in blog if kind == "open"]
    assert opens and all(r % 2 == 0 for b in opens for r in b)
    assert len(opens) == sum(1 for kind, _ in blog if kind == "close")
    assert FakeDomain.live == 0
    assert all(r.finished >= r.admitted >= r.arrival for r in res)


class DivergingDomain(FakeDomain):
    """A stand-in whose state turns non-finite (the request with rid 2)."""

    def nonfinite(self):
        return self.req.rid == 2


@pytest.mark.parametrize("mode", ["streams", "sequential", "batched"])
def test_a_diverged_request_fails_alone(mode):
    FakeDomain.live = FakeDomain.peak = 0
    reqs = _reqs(5)
    log, blog = [], []
    srv = serving.Server(10 ** 9, burst=2, domain_factory=lambda r: DivergingDomain(r, log),
                         bytes_of=lambda r: 1, clock=FakeClock(),
                         batch_factory=lambda doms: FakeBatch(doms, blog), workers=1)
    if mode == "batched":
        res = srv.run_batched(reqs)
    else:
        res = srv.run(reqs, sequential=(mode == "sequential"))
    assert [r.rid for r in res] == list(range(5))
    assert [r.ok for r in res] == [True, True, False, True, True]
    assert res[2].status.startswith("failed")
    # every request still ran all its steps and was released
    assert sorted(log) == [(r, tuple(range(1, 8))) for r in range(5)]
    assert FakeDomain.live == 0


def test_summary_leaves_failed_requests_out_of_the_rates():
    res = [serving.Result(i, 1000, 10, 0.0, 0.0, float(i + 1)) for i in range(4)]
    res.append(serving.Result(4, 10 ** 6, 10, 0.0, 0.0, 4.0, status="failed: non-finite state (diverged)"))
    s = serving.summarize(res)
    assert s["requests"] == 5 and s["failed"] == 1
    assert abs(s["mcell_updates_per_s"] - 4 * 1000 * 10 / 4.0 / 1e6) < 1e-12
    assert s["p99_s"] <= 4.0
