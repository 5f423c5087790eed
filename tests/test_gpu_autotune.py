"""The per-shape autotuner of the fused SOR iteration's launch shape and
of the stencil kernels' k levels per thread (les_domain_autotune): it
picks a valid shape, leaves the domain's state
and last press statistics untouched, is remembered per grid shape, and the
tuned domain steps bit for bit like an untuned one (every tiling gives the
colour sweeps' iterates) with the same SOR iteration counts.  The residual
history sum(reltmp^2) is summed in FP64 per SOR block, so another tiling
adds the same exact squares in another grouping: it agrees to rounding
(relative 1e-12), not bit for bit."""
import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
FIELDS = ("u", "v", "w", "p", "rhs", "sm", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "uspd",
          "vspd") + tuple(f"diu{q}" for q in range(1, 10))


@pytest.mark.parametrize("case", [dict(im=90, jm=40, km=24, seed=31), dict(im=61, jm=27, km=18, seed=32)],
                         ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}")
def test_autotune_keeps_state_and_results(case):
    s = setup_gen.build(setup_gen.small_workload(**case))
    w = s.workload
    print(f"\n[autotune] {w.im}x{w.jm}x{w.km}", flush=True)
    api.release_all()
    a, b = api.Domain(s), api.Domain(s)
    for n in (1, 2):
        a.step(n, 4)
        b.step(n, 4)
    before = {f: a.download(f) for f in FIELDS}
    st_before = a.press_stats()
    r = a.autotune(reps=4)
    assert r["rows"] in ((1, 2) if w.jm % 2 == 0 else (1,)), r
    assert r["tile_rows"] in ((16, 32) if r["rows"] == 2 else (16,)), r
    assert r["tile_width"] in ((64, 32) if r["tile_rows"] == 32 else (64,)), r
    assert 1 <= r["kchunk"] <= w.km, r
    assert r["us_per_iter"] > 0.0 and r["us_per_iter_default"] > 0.0, r
    assert r["us_per_iter"] <= r["us_per_iter_default"], r
    # the stencil kernels' k levels per thread (0: automatic), timed on the
    # domain's fields and restored (every field they write is checked below)
    for key in ("velfg_kchunk", "sm_kchunk", "visc_kchunk"):
        assert 0 <= r[key] <= w.km, r
    assert 0.0 < r["stencil_us"] <= r["stencil_us_default"], r
    for f in FIELDS:
        assert np.array_equal(a.download(f), before[f]), f
    st_after = a.press_stats()
    assert st_after["iters"] == st_before["iters"]
    assert st_after["sor"].tolist() == st_before["sor"].tolist()
    # the timed run of the chosen shape leaves the state as well
    assert a.time_sor(20) > 0.0
    for f in FIELDS:
        assert np.array_equal(a.download(f), before[f]), f
    assert a.press_stats()["sor"].tolist() == st_before["sor"].tolist()
    # remembered for the shape
    rb = b.autotune()
    assert (rb["kchunk"], rb["rows"], rb["tile_rows"], rb["tile_width"]) == (r["kchunk"], r["rows"], r["tile_rows"],
                                                                            r["tile_width"])
    c = api.Domain(s)  # an untuned twin stepping alongside
    for n in (1, 2):
        c.step(n, 4)
    for n in (3, 4):
        a.step(n, 4)
        c.step(n, 4)
    for f in FIELDS:
        assert np.array_equal(a.download(f), c.download(f)), f
    sa, sc = a.press_stats(), c.press_stats()
    assert sa["iters"] == sc["iters"]
    np.testing.assert_allclose(sa["sor"], sc["sor"], rtol=1e-12, atol=0)
    for d in (a, b, c):
        d.close()
