"""Device-side set-up and concurrent serving on the GPU.

* masks built on the device from the height map (feedbfm.f95:87-116) and
  cn1 evaluated from its separable definition (init.f95:100-107) give the
  same masks and the same steps, bit for bit, as the host-built arrays;
* a domain served together with others on concurrent streams produces the
  same fields, bit for bit, as the same domain run alone;
* a domain on device memory kept from a destroyed one steps bit for bit as
  on fresh memory.
"""
import numpy as np
import pytest

from lesamd import api, serving, setup_gen

pytestmark = pytest.mark.gpu
FIELDS = ["u", "v", "w", "p", "f", "g", "h", "usum", "sm", "fold"] + [f"diu{q}" for q in (1, 5, 9)]


@pytest.mark.parametrize("dims", [(40, 36, 20), (33, 29, 17)], ids=lambda d: "x".join(map(str, d)))
def test_device_setup_equals_host_setup(dims):
    w = setup_gen.Workload("dev-setup", *dims, seed=11)
    print(f"\n[device setup] {dims}", flush=True)
    host = api.Domain(setup_gen.build(w))
    dev = api.Domain(setup_gen.build(w, dense=False))
    for name in ("bmask1", "cmask1", "dmask1"):
        assert np.array_equal(host.download(name), dev.download(name)), name
    assert host.download("bmask1").any()
    for n in range(1, 4):
        host.step(n, 3)
        dev.step(n, 3)
    for name in FIELDS:
        assert np.array_equal(host.download(name), dev.download(name)), name
    assert host.press_stats()["iters"] == dev.press_stats()["iters"]
    host.close()
    dev.close()


def test_domain_alone_equals_domain_in_the_mix():
    reqs = serving.sample_requests(3, 6, seed=5, sizes=[(48, 40, 24), (33, 29, 17), (64, 32, 20)])
    print("\n[serving mix] 3 domains, 6 steps", flush=True)
    finals = {}

    class Keep:
        def __init__(self, req):
            self.req = req
            self.d = api.Domain(setup_gen.build(req.workload, dense=False))
            self.d.set_graph(True)

        def step(self, n, nmax):
            self.d.step(n, nmax)

        def done(self):
            return self.d.done()

        def close(self):
            finals[self.req.rid] = {k: self.d.download(k) for k in FIELDS}
            self.d.close()

    srv = serving.Server(64 << 30, burst=2, domain_factory=Keep,
                         bytes_of=lambda r: api.bytes_needed(r.workload.im, r.workload.jm, r.workload.km))
    res = srv.run(reqs)
    assert len(res) == 3
    for r in reqs:
        alone = api.Domain(setup_gen.build(r.workload, dense=False))
        alone.run(1, r.nsteps, r.nsteps)
        for k in FIELDS:
            assert np.array_equal(alone.download(k), finals[r.rid][k]), (r.rid, k)
        alone.close()


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_domain_on_reused_memory_steps_like_a_fresh_one(graph):
    """A destroyed domain's device memory is kept and handed to the next
    domain of the same extents with its last contents still in it; the new
    domain must step exactly as the first did on fresh memory (set-up clears
    what must start at zero)."""
    w = setup_gen.Workload("reuse", 48, 40, 24, seed=21)
    s = setup_gen.build(w, dense=False)
    print(f"\n[memory reuse] 48x40x24 graph={graph}", flush=True)
    api.release_all()  # no kept memory: the first domain gets fresh blocks
    runs = []
    for _ in range(2):
        d = api.Domain(s)
        d.set_graph(graph)
        for n in range(1, 6):
            d.step(n, 5)
        runs.append(({f: d.download(f) for f in FIELDS + ["rhs", "sm"]}, d.press_stats()))
        d.close()
    (a, sa), (b, sb) = runs
    for f in a:
        assert np.array_equal(a[f], b[f]), f
    assert sa["iters"] == sb["iters"]
    assert sa["sor"].tolist() == sb["sor"].tolist()
