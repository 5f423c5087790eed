"""Batched steps (les_batch_*): domains advanced together, their SOR
iterations in one launch per iteration, give each domain exactly the
results of stepping it alone -- the same result alone and in the mix --
with eager steps, graph replays, a domain whose SOR cannot be batched (odd
jm) in the mix, and reads through the domain API between steps."""
import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
CASES = [dict(im=40, jm=30, km=20, seed=41), dict(im=33, jm=26, km=17, seed=42), dict(im=24, jm=21, km=12, seed=43),
         dict(im=70, jm=16, km=9, seed=44)]
FIELDS = ("u", "v", "w", "p", "rhs", "sm", "f", "g", "h", "usum", "fold")
NSTEPS = 5


def test_batch_equals_alone():
    setups = [setup_gen.build(setup_gen.small_workload(**c)) for c in CASES]
    print(f"\n[batch] {len(setups)} domains, {NSTEPS} steps", flush=True)
    api.release_all()
    alone = []
    for s in setups:
        d = api.Domain(s)
        d.set_graph(True)
        for n in range(1, NSTEPS + 1):
            d.step(n, NSTEPS)
        alone.append(({f: d.download(f) for f in FIELDS}, d.press_stats()))
        d.close()
    doms = [api.Domain(s) for s in setups]
    b = api.Batch(doms)
    mid = None
    for n in range(1, NSTEPS + 1):
        b.step([n] * len(doms), [NSTEPS] * len(doms))
        if n == 3:
            mid = doms[0].download("p")  # ordered after the batch's work on this domain
    b.synchronize()
    for q, d in enumerate(doms):
        want, st = alone[q]
        for f in FIELDS:
            assert np.array_equal(d.download(f), want[f]), (q, f)
        got = d.press_stats()
        assert got["iters"] == st["iters"], q
        assert got["sor"].tolist() == st["sor"].tolist(), q
    b.close()
    # the mid-run read saw step 3 exactly
    d = api.Domain(setups[0])
    for n in range(1, 4):
        d.step(n, NSTEPS)
    assert np.array_equal(mid, d.download("p"))
    d.close()
    for d in doms:
        d.close()
