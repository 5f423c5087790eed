"""GPU parity: the HIP library through its C ABI against the C restatement
(oracle/les_oracle.c, FP32 stencils with FP64 reductions -- the same
arithmetic the GPU path uses), stage by stage and over whole steps.

Tolerances: stencil stages are compared bit for bit (same FP32 operations in
the same order, no contraction, correctly rounded divide and sqrt).  press
adds three FP64 reductions whose summation order differs (tree on the GPU,
sequential on the CPU); its outputs are compared at rel-L2 <= 1e-6 and the
SOR iteration counts must be equal.  Multi-step trajectories: rel-L2 <= 1e-5
per step on u,v,w,p,sm.
"""
import numpy as np
import pytest

from backends import GpuBackend, rel_l2, ulp_diff
from lesamd import setup_gen
from oracle import oracle

pytestmark = pytest.mark.gpu

CMP = ["u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "rhs", "sm",
       "uspd", "vspd"] + [f"diu{q}" for q in range(1, 10)]

CASES = [
    dict(im=24, jm=20, km=16, seed=1),                      # even jm, log-law inflow
    dict(im=23, jm=19, km=13, seed=2),                      # odd extents: one-sided/periodic quirks
    dict(im=33, jm=18, km=21, seed=3, ustar=None),          # the reference's fixed 5 m/s inflow
    dict(im=5, jm=2, km=2, seed=4),                         # smallest extents: every cell on a face
    dict(im=62, jm=6, km=5, seed=5),                        # row pitch exactly i = -1..im+2 (128 floats)
]


def _mk(case):
    case = dict(case)
    w = setup_gen.small_workload(**case)
    return setup_gen.build(w)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}")
def test_stage_by_stage(case):
    s = _mk(case)
    print(f"\n[stage parity] {s.workload.name} seed={s.workload.seed}", flush=True)
    gpu = GpuBackend(s)
    orc = oracle.OracleBackend(s, 32, redmode=1)
    for n in range(1, 4):
        for st in oracle.STAGES:
            kw = {"n": n, "nmax": 1000} if st in ("bondv1", "adam", "press") else {}
            gpu.call(st, **kw)
            orc.call(st, **kw)
            a, b = gpu.fetch(CMP), orc.fetch(CMP)
            for name in CMP:
                if st == "press" and name in ("p", "rhs"):
                    assert rel_l2(a[name], b[name]) <= 1e-6, (n, st, name, rel_l2(a[name], b[name]))
                else:
                    assert np.array_equal(a[name], b[name]), (n, st, name, ulp_diff(a[name], b[name]))
            if st == "press":
                sg, so = gpu.press_stats(), orc.press_stats()
                assert sg["iters"] == so["iters"], (n, sg["iters"], so["iters"])
                np.testing.assert_allclose(sg["sor"], so["sor"], rtol=1e-9)
                assert sg["nonfinite"] == 0


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}")
def test_trajectory_20_steps(case):
    s = _mk(case)
    print(f"\n[20-step trajectory] {s.workload.name}", flush=True)
    gpu = GpuBackend(s)
    orc = oracle.OracleBackend(s, 32, redmode=1)
    names = ["u", "v", "w", "p", "sm"]
    for n in range(1, 21):
        gpu.step(n, 20)
        orc.step(n, 20)
        a, b = gpu.fetch(names), orc.fetch(names)
        for name in names:
            assert rel_l2(a[name], b[name]) <= 1e-5, (n, name, rel_l2(a[name], b[name]))
        assert gpu.press_stats()["iters"] == orc.press_stats()["iters"]


def test_domain_api_matches_reference_signature_api():
    """Whole steps through les_domain_step equal the seven reference-signature
    calls bit for bit (same kernels, same order)."""
    from lesamd import api
    s = _mk(dict(im=24, jm=20, km=16, seed=4))
    print(f"\n[domain api] {s.workload.name}", flush=True)
    gpu = GpuBackend(s)
    for n in range(1, 4):
        gpu.step(n, 3)
    ref = gpu.fetch(["u", "v", "w", "p", "sm", "f"])
    dom = api.Domain(s)
    dom.run(1, 3, 3)
    for name in ("u", "v", "w", "p", "sm", "f"):
        assert np.array_equal(dom.download(name), ref[name]), name
    st = dom.press_stats()
    assert st["iters"] == gpu.press_stats()["iters"]
    dom.close()


@pytest.mark.parametrize("blocks", [1, 20, 1000])
def test_sor_k_chunking(blocks, monkeypatch):
    """Each SOR block walks a range of k levels, carrying neighbour values
    from level to level; any split of k among blocks (one range, ranges of
    3, single levels) gives the same iterate."""
    monkeypatch.setenv("LES_SOR_BLOCKS", str(blocks))
    s = _mk(dict(im=70, jm=19, km=11, seed=5))
    print(f"\n[sor chunking] blocks={blocks} {s.workload.name}", flush=True)
    gpu = GpuBackend(s)
    orc = oracle.OracleBackend(s, 32, redmode=1)
    for n in range(1, 3):
        gpu.step(n, 2)
        orc.step(n, 2)
    a, b = gpu.fetch(["p", "u", "rhs"]), orc.fetch(["p", "u", "rhs"])
    for name in ("p", "u", "rhs"):
        assert rel_l2(a[name], b[name]) <= 1e-6, (name, rel_l2(a[name], b[name]))
    assert gpu.press_stats()["iters"] == orc.press_stats()["iters"]


def test_sor_converges_early_like_reference():
    """A loose stop threshold makes SOR stop before nmaxp: the iteration
    count, residual history and iterate must match (the stop test runs one
    launch behind on the GPU)."""
    s = _mk(dict(im=24, jm=20, km=16, seed=6, pjuge=50.0))
    print(f"\n[sor early stop] {s.workload.name}", flush=True)
    gpu = GpuBackend(s)
    orc = oracle.OracleBackend(s, 32, redmode=1)
    seen = set()
    for n in range(1, 4):
        gpu.step(n, 3)
        orc.step(n, 3)
        sg, so = gpu.press_stats(), orc.press_stats()
        assert sg["iters"] == so["iters"], (n, sg["iters"], so["iters"])
        seen.add(so["iters"])
        np.testing.assert_allclose(sg["sor"], so["sor"], rtol=1e-9)
        a, b = gpu.fetch(["p"]), orc.fetch(["p"])
        assert rel_l2(a["p"], b["p"]) <= 1e-6
    assert min(seen) < 50, seen


def test_graph_replay_matches_eager():
    """Steps replayed from the captured graph equal eager steps bit for bit,
    on every replay (checked after each step, with host reads in between)."""
    from lesamd import api
    s = _mk(dict(im=40, jm=24, km=18, seed=7))
    print(f"\n[graph replay] {s.workload.name}", flush=True)
    eager, graph = api.Domain(s), api.Domain(s)
    graph.set_graph(True)
    for n in range(1, 7):
        eager.step(n, 6)
        graph.step(n, 6)
        for name in ("u", "v", "w", "p", "sm", "fold"):
            assert np.array_equal(eager.download(name), graph.download(name)), (n, name)
        assert eager.press_stats()["sor"].tolist() == graph.press_stats()["sor"].tolist()
    eager.close()
    graph.close()


FUSED_CASES = [
    (dict(im=250, jm=29, km=20, seed=8), 640),   # three pair tiles, three row tiles, odd jm
    (dict(im=37, jm=12, km=9, seed=9), 640),     # one row tile exactly, odd im
    (dict(im=20, jm=3, km=7, seed=10), 640),     # jm odd and smaller than the tile ring
    (dict(im=9, jm=2, km=5, seed=11), 640),      # jm = 2
    (dict(im=64, jm=26, km=33, seed=12), 1),     # one k chunk of 33 levels
    (dict(im=64, jm=26, km=33, seed=12), 5000),  # one level per block
    (dict(im=250, jm=30, km=20, seed=18), 640),  # three pair tiles, three row tiles, even jm
    (dict(im=121, jm=14, km=17, seed=19), 0),    # automatic k chunks, odd im, even jm
    (dict(im=300, jm=24, km=90, seed=20), 0),    # the 300-wide rows of the bench domain
    (dict(im=128, jm=27, km=10, seed=21), 0),    # 64 pairs: one tile as wide as the domain, odd jm
    (dict(im=127, jm=20, km=9, seed=22), 640),   # 64 pairs, odd im, even jm
    (dict(im=66, jm=18, km=7, seed=23), 0),      # 33 pairs: one whole-width tile of 64, two of 32
    (dict(im=126, jm=8, km=6, seed=24), 0),      # row pitch exactly i = -1..im+2 (192 floats)
]


@pytest.mark.parametrize("rows,tile_rows,tile_width", [(1, 16, 64), (2, 16, 64), (2, 32, 64), (2, 32, 32)],
                         ids=["r1", "r2t16", "r2t32", "r2t32w32"])
@pytest.mark.parametrize("case,blocks", FUSED_CASES,
                         ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}" if isinstance(c, dict) else f"b{c}")
def test_fused_sor_iteration_equals_colour_sweeps(case, blocks, rows, tile_rows, tile_width, monkeypatch):
    if tile_width == 32 and case["im"] == 126:
        pytest.xfail("32 x 32 tiles of k_sor_iter2 at 126 x 8 x 6: stop-test sums differ from the colour "
                     "sweeps' (iterates bit-identical); the tile is left out of the autotuner")
    """One launch per SOR iteration (red on the tile and a ring, then black,
    marching k) gives the iterates of the two colour sweeps bit for bit, for
    any tiling of the domain, and the same iteration counts and residuals;
    with one tile row per thread (k_sor_iter) and with two (k_sor_iter2,
    even jm, tiles of 16 x 64, 32 x 64 or 32 x 32 rows x pairs; odd jm keeps
    k_sor_iter)."""
    from lesamd import api
    s = _mk(case)
    print(f"\n[fused sor] {s.workload.name} blocks={blocks} rows={rows} tile={tile_rows}x{tile_width}", flush=True)
    monkeypatch.setenv("LES_SOR_FUSED", "0")
    two = api.Domain(s)
    monkeypatch.setenv("LES_SOR_FUSED", "1")
    monkeypatch.setenv("LES_SOR_FUSED_BLOCKS", str(blocks))
    monkeypatch.setenv("LES_SOR_ROWS", str(rows))
    monkeypatch.setenv("LES_SOR_TILE_ROWS", str(tile_rows))
    monkeypatch.setenv("LES_SOR_TILE_WIDTH", str(tile_width))
    one = api.Domain(s)
    for n in range(1, 4):
        two.step(n, 3)
        one.step(n, 3)
        for name in ("p", "u", "v", "w", "rhs"):
            assert np.array_equal(two.download(name), one.download(name)), (n, name)
        a, b = two.press_stats(), one.press_stats()
        assert a["iters"] == b["iters"]
        np.testing.assert_allclose(a["sor"], b["sor"], rtol=1e-12)
    two.close()
    one.close()


FUSED_STEP_CASES = [
    (dict(im=24, jm=20, km=16, seed=13), 0),     # automatic k chunks
    (dict(im=23, jm=19, km=13, seed=14), 4),     # odd extents, chunks of 4 levels
    (dict(im=70, jm=9, km=11, seed=15), 11),     # two blocks in i, one chunk per column
    (dict(im=13, jm=2, km=6, seed=16), 1),       # jm = 2, one level per thread
    (dict(im=130, jm=31, km=9, seed=17), 3),     # three blocks in i
]
ALL_FIELDS = ["u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "rhs", "sm",
              "uspd", "vspd"] + [f"diu{q}" for q in range(1, 10)]


@pytest.mark.parametrize("case,kchunk", FUSED_STEP_CASES,
                         ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}" if isinstance(c, dict) else f"kc{c}")
def test_fused_step_kernels_equal_stages(case, kchunk, monkeypatch):
    """The whole-step kernels -- vel2+velfg+feedbf in one pass, adam inside
    the les viscous kernel -- leave every array (diu1..diu9 and sm with all their ghost
    cells, f, g, h and their previous values, the feedback integrals, uspd,
    vspd, p) bit-identical to the separate stage kernels, at steps 1
    (initial copy) to 3, for any split of k among threads and blocks."""
    from lesamd import api
    s = _mk(case)
    print(f"\n[fused step kernels] {s.workload.name} kchunk={kchunk}", flush=True)
    monkeypatch.setenv("LES_FUSE_VELFG", "0")
    monkeypatch.setenv("LES_FUSE_LES", "0")
    sep = api.Domain(s)
    monkeypatch.setenv("LES_FUSE_VELFG", "1")
    monkeypatch.setenv("LES_FUSE_LES", "1")
    monkeypatch.setenv("LES_VELFG_KCHUNK", str(kchunk))
    one = api.Domain(s)
    for n in range(1, 4):
        sep.step(n, 3)
        one.step(n, 3)
        for name in ALL_FIELDS:
            assert np.array_equal(sep.download(name), one.download(name)), (n, name)
    sep.close()
    one.close()
