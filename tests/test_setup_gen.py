"""Set-up generator: metrics, Poisson coefficients and run constants against
the reference's own grid/init/set subroutines; city, masks and inflow
against their closed forms."""
import math

import numpy as np
import pytest

from lesamd import layout, setup_gen
from oracle import oracle


def test_reference_grid_rule_equals_grid_f95(ref_small):
    """grid.f95:36-172 at the build's extents (25x25x105)."""
    w = setup_gen.Workload("g", 25, 25, 105, dz_rule="reference", city=None, ustar=None)
    m = setup_gen.grid_metrics(w)
    ref = oracle.RefBackend(setup_gen.build(w), ref_small).ref_grid()
    for name in ("dx1", "dxl", "dxs", "dy1", "dyl", "dys", "dzn", "dzs", "z2"):
        np.testing.assert_array_equal(m[name], ref[name], err_msg=name)


@pytest.mark.parametrize("dz_rule", ["reference", "stretched"])
def test_poisson_coefficients_equal_init_f95(dz_rule, ref_small):
    """init.f95:82-107 on the generator's metrics; init also zeroes u,v,w,p."""
    w = setup_gen.Workload("c", 21, 17, 40, dz_rule=dz_rule, city=None, ustar=None)
    s = setup_gen.build(w)
    out = oracle.RefBackend(s, ref_small).ref_init(s.metrics)
    for name in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s", "cn1"):
        np.testing.assert_array_equal(s.coeffs[name], out[name], err_msg=name)
    for name in ("u", "v", "w", "p"):
        assert not out[name].any(), name


def test_run_constants_equal_set_f95(ref_small):
    w = setup_gen.small_workload(ustar=None)
    c = oracle.RefBackend(setup_gen.build(w), ref_small).ref_set()
    assert np.float32(c["ro"]) == setup_gen.RO
    assert np.float32(c["vn"]) == setup_gen.VN
    assert np.float32(c["alpha"]) == setup_gen.ALPHA and np.float32(c["beta"]) == setup_gen.BETA
    assert np.float32(c["dt"]) == np.float32(0.05)


def test_stretched_dz_rule():
    """dz1=2 m, x1.05 per level, capped at 20 m; dzs the midpoint spacing."""
    w = setup_gen.Workload("s", 8, 8, 90, city=None, ustar=None)
    m = setup_gen.grid_metrics(w)
    dzn = m["dzn"][2: w.km + 2]  # k = 1..km
    assert dzn[0] == np.float32(2.0)
    grow = dzn[1:] / dzn[:-1]
    capped = dzn[1:] == np.float32(20.0)
    assert np.all(np.abs(grow[~capped] - 1.05) < 1e-6)
    assert dzn.max() == np.float32(20.0)
    assert capped.sum() > 0
    dzs = m["dzs"][2: w.km + 1]
    np.testing.assert_allclose(dzs, 0.5 * (dzn[:-1] + dzn[1:]), rtol=1e-6)


def test_city_geometry_and_heights():
    """Blocks of 8x8 cells separated by 2-cell streets, 15x15 blocks on
    150x150, lognormal heights of median ~16 m."""
    w = setup_gen.workloads()["plain-150x150x90"]
    zb = setup_gen.city_heights(w)
    assert zb.shape == (150, 150)
    built = zb > 0
    assert built.sum() == 15 * 15 * 64
    # each block is constant and 8 wide, streets 2 wide
    row = built[:, 5]
    runs = np.diff(np.flatnonzero(np.diff(np.r_[0, row.astype(int), 0])))
    assert set(runs[::2]) == {8}
    heights = zb[1::10, 1::10]
    assert heights.shape == (15, 15)
    med = float(np.median(heights))
    assert 10.0 < med < 26.0
    assert np.all(zb[1:9, 1:9] == zb[1, 1])


def test_mask_rule_known_answer():
    """feedbfm.f95:87-116: a cell is solid where zbm > z2(k)+0.5*dzn(k); with
    dz 2 m uniform (z2(1)=2, z2(k)=2k) the levels below a 7 m building are
    k=1 (3 m), k=2 (5 m), k=3 (7 m is not < 7)."""
    w = setup_gen.Workload("m", 4, 4, 6, dz1=2.0, dz_growth=1.0, city=None, ustar=None)
    m = setup_gen.grid_metrics(w)
    zb = np.zeros((4, 4), np.float32)
    zb[1, 2] = 7.0
    masks = setup_gen.building_masks(zb, m, 4, 4, 6)
    a = layout.interior("amask1", masks["amask1"], 4, 4, 6)
    b = layout.interior("bmask1", masks["bmask1"], 4, 4, 6)
    assert a[1, 2].tolist() == [0, 0, 1, 1, 1, 1]
    assert (b + a == 1).all()
    assert a.sum() == 4 * 4 * 6 - 2
    # ghosts stay zero
    assert masks["bmask1"].sum() == b.sum() and masks["dmask1"].sum() == b.sum()


def test_loglaw_inflow():
    w = setup_gen.small_workload(8, 6, 12, ustar=0.5, z0=0.1, noise=0.0)
    m = setup_gen.grid_metrics(w)
    prof = setup_gen.inflow_profile(w, m)
    zc = np.array([m["z2"][k] + 0.5 * m["dzn"][k + 1] for k in range(1, 13)], dtype=np.float64)
    want = (0.5 / 0.4) * np.log(zc / 0.1)
    np.testing.assert_allclose(prof[0], want.astype(np.float32), rtol=1e-6)
    assert np.all(prof == prof[0:1])
    w2 = setup_gen.small_workload(8, 6, 12, ustar=0.5, z0=0.1, noise=0.01, seed=3)
    p2 = setup_gen.inflow_profile(w2, m)
    r = p2 / want[None, :]
    assert np.all(np.abs(r - 1) <= 0.0100001) and np.abs(r - 1).max() > 0.005


def test_generator_is_deterministic():
    a = setup_gen.build(setup_gen.small_workload(seed=9))
    b = setup_gen.build(setup_gen.small_workload(seed=9))
    c = setup_gen.build(setup_gen.small_workload(seed=10))
    assert np.array_equal(a.zbm, b.zbm) and np.array_equal(a.inflow, b.inflow)
    assert not np.array_equal(a.inflow, c.inflow)
