"""Two SOR iterations per launch (k_sor_pair, LES_SOR_ROWS=4, the default)
against one iteration per launch (k_sor_iter2, LES_SOR_ROWS=2), whose
iterates equal the colour sweeps and the C restatement bit for bit
(tests/test_gpu_parity.py): fields bit for bit, the same iteration counts,
the FP64 residual history to rounding (the two kernels group the squares in
other blocks), over tilings and k chunks, odd and even iteration caps,
stops at the first and at the second iteration of a launch (the first needs
the fix-up launch), and the entry ghosts of p perturbed by an upload (the
launch of iterations 1 and 2 then reads them)."""
import dataclasses
import os

import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
FIELDS = ("p", "u", "v", "w", "rhs", "sm")


def _domain(s, **env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return api.Domain(s)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _compare(ref, pair, n):
    for f in FIELDS:
        assert np.array_equal(ref.download(f), pair.download(f)), (n, f)
    a, b = ref.press_stats(), pair.press_stats()
    assert a["iters"] == b["iters"], (n, a["iters"], b["iters"])
    np.testing.assert_allclose(b["sor"], a["sor"], rtol=1e-12)
    return a


CASES = [
    (dict(im=24, jm=20, km=16, seed=31), 0, 50),    # automatic k chunk
    (dict(im=121, jm=14, km=17, seed=32), 3, 50),   # odd im, chunks of 3 levels
    (dict(im=64, jm=26, km=33, seed=33), 33, 7),    # one chunk, odd iteration cap
    (dict(im=250, jm=30, km=20, seed=34), 5, 50),   # five pair tiles x two row tiles
    (dict(im=9, jm=2, km=5, seed=35), 1, 9),        # jm = 2, one level per block, odd cap
    (dict(im=300, jm=24, km=90, seed=36), 0, 50),   # 300-wide rows of the benchmark domain
    (dict(im=130, jm=50, km=12, seed=37), 4, 1),    # a cap of one iteration
    (dict(im=120, jm=48, km=9, seed=38), 2, 4),     # 60 pairs: one tile as wide as the domain
    (dict(im=128, jm=20, km=12, seed=39), 0, 50),   # 64 pairs: the widest single tile
    (dict(im=129, jm=22, km=9, seed=40), 3, 7),     # 65 pairs: two tiles with their rings
]


@pytest.mark.parametrize("case,kchunk,nmaxp", CASES,
                         ids=[f"{c['im']}x{c['jm']}x{c['km']}-kc{k}-n{n}" for c, k, n in CASES])
def test_pair_equals_one_iteration_kernel(case, kchunk, nmaxp):
    w = dataclasses.replace(setup_gen.small_workload(**case), nmaxp=nmaxp)
    s = setup_gen.build(w, dense=False)
    print(f"\n[sor pair] {w.name} kchunk={kchunk} nmaxp={nmaxp}", flush=True)
    ref = _domain(s, LES_SOR_ROWS=2)
    pair = _domain(s, LES_SOR_ROWS=4, LES_SOR_KCHUNK=kchunk)
    for n in range(1, 4):
        ref.step(n, 3)
        pair.step(n, 3)
        _compare(ref, pair, n)
    ref.close()
    pair.close()


@pytest.mark.parametrize("stop_at", [1, 2, 3, 4, 7, 12])
def test_pair_stops_where_one_iteration_stops(stop_at):
    """pjuge between the residuals of iterations stop_at - 1 and stop_at of
    step 1 (from a run that never stops): both kernels stop at stop_at, odd
    (the first of a launch: its result comes from the fix-up) or even."""
    base = setup_gen.small_workload(im=40, jm=28, km=14, seed=41)
    probe = _domain(setup_gen.build(dataclasses.replace(base, pjuge=1e-30), dense=False), LES_SOR_ROWS=2)
    probe.step(1, 2)
    h = probe.press_stats()["sor"]
    probe.close()
    lo = h[stop_at - 1]
    hi = min(h[: stop_at - 1]) if stop_at > 1 else np.inf
    assert lo < hi, (stop_at, h[: stop_at + 1])
    pj = float(np.sqrt(lo * hi)) if np.isfinite(hi) else float(lo * 2)
    w = dataclasses.replace(base, pjuge=pj)
    s = setup_gen.build(w, dense=False)
    print(f"\n[sor pair stop] at iteration {stop_at}, pjuge {pj:.4e}", flush=True)
    ref = _domain(s, LES_SOR_ROWS=2)
    pair = _domain(s, LES_SOR_ROWS=4)
    for n in (1, 2):
        ref.step(n, 2)
        pair.step(n, 2)
        st = _compare(ref, pair, n)
        if n == 1:
            assert st["iters"] == stop_at
    ref.close()
    pair.close()


@pytest.mark.parametrize("face", ["all", "i0", "jm1", "k0", "km1"])
def test_pair_reads_the_entry_ghosts(face):
    """An upload of p whose ghosts differ from their boundp copies: the
    launch of iterations 1 and 2 reads the entry values as the reference's
    first sweep does, the same bits as the entry-ghost kernel."""
    w = setup_gen.Workload("sor-pair-entry", 64, 48, 20, seed=43)
    s = setup_gen.build(w, dense=False)
    print(f"\n[sor pair entry ghosts] {face}", flush=True)
    ref = _domain(s, LES_SOR_ROWS=2)
    pair = _domain(s, LES_SOR_ROWS=4)
    for n in (1, 2):
        ref.step(n, 6)
        pair.step(n, 6)
    p = ref.download("p")
    im, jm, km = w.im, w.jm, w.km
    faces = {"i0": np.s_[0, 1:jm + 1, 1:km + 1], "im1": np.s_[im + 1, 1:jm + 1, 1:km + 1],
             "j0": np.s_[1:im + 1, 0, 1:km + 1], "jm1": np.s_[1:im + 1, jm + 1, 1:km + 1],
             "k0": np.s_[1:im + 1, 1:jm + 1, 0], "km1": np.s_[1:im + 1, 1:jm + 1, km + 1]}
    rng = np.random.default_rng(8)
    for f in (faces if face == "all" else [face]):
        p[faces[f]] += rng.uniform(-0.5, 0.5, p[faces[f]].shape).astype(np.float32)
    ref.upload("p", p)
    pair.upload("p", p)
    for n in (3, 4):
        ref.step(n, 6)
        pair.step(n, 6)
        _compare(ref, pair, n)
    ref.close()
    pair.close()


def test_pair_graph_replay_and_autotune():
    """Graph replay of the pair path equals eager launches bit for bit, and
    the autotuner's choice (it times the pair kernel among the candidates)
    leaves the results unchanged."""
    w = setup_gen.Workload("sor-pair-graph", 90, 40, 24, seed=44)
    s = setup_gen.build(w, dense=False)
    print("\n[sor pair graph + autotune] 90x40x24", flush=True)
    eager = _domain(s, LES_SOR_ROWS=4)
    graph = _domain(s, LES_SOR_ROWS=4)
    graph.set_graph(True)
    tuned = _domain(s)
    t = tuned.autotune(reps=4)
    print(f"  autotune: {t}", flush=True)
    for n in range(1, 6):
        for d in (eager, graph, tuned):
            d.step(n, 5)
        for f in FIELDS:
            a = eager.download(f)
            assert np.array_equal(a, graph.download(f)), (n, f)
            assert np.array_equal(a, tuned.download(f)), (n, f)
    for d in (eager, graph, tuned):
        d.close()


def test_pair_head_rhs_downloads():
    """The one-pass head leaves the reference-layout rhs to a gather on
    download: rhs equals the one-iteration path's after eager steps, after
    graph replays with downloads in between, after an upload (which the
    download returns as uploaded), and after a press of mode 1, whose head
    writes rhs itself."""
    w = setup_gen.Workload("sor-pair-head", 70, 36, 22, seed=45)
    s = setup_gen.build(w, dense=False)
    print("\n[sor pair head] rhs downloads 70x36x22", flush=True)
    ref = _domain(s, LES_SOR_ROWS=2)
    pair = _domain(s, LES_SOR_ROWS=4)
    for n in (1, 2):
        ref.step(n, 8)
        pair.step(n, 8)
    _compare(ref, pair, 2)
    for d in (ref, pair):
        d.set_graph(True)
    for n in (3, 4, 5):
        ref.step(n, 8)
        pair.step(n, 8)
        if n != 4:
            _compare(ref, pair, n)
    _compare(ref, pair, 5)
    r = np.random.default_rng(3).uniform(-1, 1, ref.download("rhs").shape).astype(np.float32)
    for d in (ref, pair):  # both: press leaves the ghosts of rhs as they are
        d.upload("rhs", r)
    assert np.array_equal(pair.download("rhs"), r)
    for d in (ref, pair):
        d.set_press_mode(1, outer=2, inner=5)
    ref.step(6, 8)
    pair.step(6, 8)
    for f in FIELDS:
        assert np.array_equal(ref.download(f), pair.download(f)), f
    ref.close()
    pair.close()
