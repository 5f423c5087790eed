"""Batched steps (les_batch_*): domains advanced together -- every stage of
a step one launch for all of them, the SOR iterations one per launch (two
per launch with LES_BATCH_SOR_PAIRS=1) -- give each domain exactly the results of stepping it alone (the
same result alone and in the mix), with eager steps and graph replays, with
domains that step on their own streams in the mix (odd jm: SOR not
batchable; les and adam unfused: stages not batched), with domains joining
at different steps (a first step is not stage-batched), and with reads
through the domain API between steps."""
import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
CASES = [dict(im=40, jm=30, km=20, seed=41), dict(im=33, jm=26, km=17, seed=42), dict(im=24, jm=21, km=12, seed=43),
         dict(im=70, jm=16, km=9, seed=44), dict(im=36, jm=22, km=14, seed=45)]
UNFUSED = 4  # CASES[4] is created with les and adam as separate stages
FIELDS = ("u", "v", "w", "p", "rhs", "sm", "f", "g", "h", "usum", "fold")
NSTEPS = 5


def _domain(q, s, monkeypatch):
    if q == UNFUSED:
        monkeypatch.setenv("LES_FUSE_LES", "0")
    d = api.Domain(s)
    monkeypatch.delenv("LES_FUSE_LES", raising=False)
    return d


def _alone(setups, nsteps, monkeypatch):
    out = []
    for q, s in enumerate(setups):
        d = _domain(q, s, monkeypatch)
        d.set_graph(True)
        for n in range(1, nsteps[q] + 1):
            d.step(n, nsteps[q])
        out.append(({f: d.download(f) for f in FIELDS}, d.press_stats()))
        d.close()
    return out


def _check(doms, alone):
    """Fields bit for bit and SOR iteration counts equal.  The residual
    history sum(reltmp^2) is summed in FP64 per SOR block: a batch runs its
    own k chunk, so the same exact squares are added in another grouping and
    the history agrees to rounding (relative 1e-12), not bit for bit."""
    for q, d in enumerate(doms):
        want, st = alone[q]
        for f in FIELDS:
            assert np.array_equal(d.download(f), want[f]), (q, f)
        got = d.press_stats()
        assert got["iters"] == st["iters"], q
        np.testing.assert_allclose(got["sor"], st["sor"], rtol=1e-12, atol=0, err_msg=str(q))


@pytest.mark.parametrize("pairs", [True, False])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_alone(graph, pairs, monkeypatch):
    setups = [setup_gen.build(setup_gen.small_workload(**c)) for c in CASES]
    print(f"\n[batch] {len(setups)} domains, {NSTEPS} steps, graph={graph}, pairs={pairs}", flush=True)
    api.release_all()
    alone = _alone(setups, [NSTEPS] * len(setups), monkeypatch)
    doms = [_domain(q, s, monkeypatch) for q, s in enumerate(setups)]
    monkeypatch.setenv("LES_BATCH_SOR_PAIRS", "1" if pairs else "0")
    b = api.Batch(doms, graph=graph)
    mid = None
    for n in range(1, NSTEPS + 1):
        b.step([n] * len(doms), [NSTEPS] * len(doms))
        if n == 3:
            mid = doms[0].download("p")  # ordered after the batch's work on this domain
    b.synchronize()
    _check(doms, alone)
    b.close()
    # the mid-run read saw step 3 exactly
    d = api.Domain(setups[0])
    for n in range(1, 4):
        d.step(n, NSTEPS)
    assert np.array_equal(mid, d.download("p"))
    d.close()
    for d in doms:
        d.close()


def test_batch_staggered_joins(monkeypatch):
    """Domain q joins at step 1 + q of the batch's clock and runs NSTEPS + 2 - q
    steps: steps with a newcomer run its first step on its own stream, the
    later ones (clock 5..7, a graph capture and two replays) stage-batched;
    the batch is rebuilt whenever its members change."""
    setups = [setup_gen.build(setup_gen.small_workload(**c)) for c in CASES[:4]]
    nsteps = [NSTEPS + 2 - q for q in range(len(setups))]
    print(f"\n[batch] staggered joins, steps {nsteps}", flush=True)
    api.release_all()
    alone = _alone(setups, nsteps, monkeypatch)
    doms = [_domain(q, s, monkeypatch) for q, s in enumerate(setups)]
    b, members = None, None
    for t in range(1, NSTEPS + 3):
        act = [q for q in range(len(doms)) if 1 + q <= t <= q + nsteps[q]]
        if act != members:
            if b is not None:
                b.close()
            b = api.Batch([doms[q] for q in act])
            members = act
        b.step([t - q for q in act], [nsteps[q] for q in act])
    b.synchronize()
    b.close()
    _check(doms, alone)
    for d in doms:
        d.close()


def test_batch_pairs_stop_and_cap_per_domain(monkeypatch):
    """The batched SOR runs iterations 2.. two per launch over the shared
    launch sequence: each domain runs up to its own iteration cap (1, 4, 7
    and 50: none, whole launches, a lone last iteration) and stops at its own
    iteration (1, the first of the sequence's launches, the second of one,
    and a later first one, which need the fix-up), with the results of the
    domain stepped alone."""
    import dataclasses
    shapes = [dict(im=40, jm=28, km=14, seed=51 + q) for q in range(8)]
    # thresholds between two residuals of a run that never stops, so that
    # step 1 stops exactly at the given iteration
    stops = {4: 1, 5: 2, 6: 3, 7: 6}
    setups = []
    for q, c in enumerate(shapes):
        w = setup_gen.small_workload(**c)
        if q < 4:
            w = dataclasses.replace(w, nmaxp=(1, 4, 7, 50)[q], pjuge=1e-30)
        else:
            probe = api.Domain(setup_gen.build(dataclasses.replace(w, pjuge=1e-30)))
            probe.step(1, 3)
            h = probe.press_stats()["sor"]
            probe.close()
            at = stops[q]
            lo, hi = h[at - 1], (min(h[: at - 1]) if at > 1 else np.inf)
            assert lo < hi, (q, h[: at + 1])
            w = dataclasses.replace(w, pjuge=float(np.sqrt(lo * hi)) if np.isfinite(hi) else float(lo * 2))
        setups.append(setup_gen.build(w))
    print("\n[batch pairs] caps 1/4/7/50, stops at 1/2/3/6", flush=True)
    api.release_all()
    alone = _alone(setups, [3] * len(setups), monkeypatch)
    doms = [api.Domain(s) for s in setups]
    monkeypatch.setenv("LES_BATCH_SOR_PAIRS", "1")
    b = api.Batch(doms, graph=True)
    for n in range(1, 4):
        b.step([n] * len(doms), [3] * len(doms))
        if n == 1:
            b.synchronize()
            for q in range(4, 8):
                assert doms[q].press_stats()["iters"] == stops[q], (q, doms[q].press_stats()["iters"])
    b.synchronize()
    _check(doms, alone)
    b.close()
    for d in doms:
        d.close()
