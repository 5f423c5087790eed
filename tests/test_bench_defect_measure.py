"""bench.py's defect measure sum((cn1*d)^2), d = rhs - B(p), on which the
precision comparison ranks every press, against a cell-by-cell loop over
the pressure operator of press.f95:95-98 on a small grid (no GPU)."""
import importlib.util
from pathlib import Path

import numpy as np

from lesamd import layout, setup_gen

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _loop_measure(s, p, rhs):
    w = s.workload
    c = s.coeffs
    ax, by, cz = c["_cn1_terms"]
    tot = 0.0
    for k in range(1, w.km + 1):
        for j in range(1, w.jm + 1):
            for i in range(1, w.im + 1):
                cn = float(np.float32(1.0) / np.float32(np.float32(ax[i - 1] + by[j - 1]) + cz[k - 1]))
                b = (float(c["cn2l"][i - 1]) * p[i + 1, j, k] + float(c["cn2s"][i - 1]) * p[i - 1, j, k]
                     + float(c["cn3l"][j - 1]) * p[i, j + 1, k] + float(c["cn3s"][j - 1]) * p[i, j - 1, k]
                     + float(c["cn4l"][k - 1]) * p[i, j, k + 1] + float(c["cn4s"][k - 1]) * p[i, j, k - 1]
                     - p[i, j, k] / cn)
                tot += (cn * (rhs[i, j, k] - b)) ** 2
    return tot


def test_defect_measure_equals_the_operator_loop():
    s = setup_gen.build(setup_gen.small_workload(im=7, jm=6, km=5, seed=3), dense=False)
    w = s.workload
    rng = np.random.default_rng(0)
    p = layout.new_field("p", w.im, w.jm, w.km)
    p[...] = rng.uniform(-1, 1, p.shape).astype(np.float32)
    rhs = layout.new_field("rhs", w.im, w.jm, w.km)
    rhs[...] = rng.uniform(-1, 1, rhs.shape).astype(np.float32)
    got = _bench().defect_measure(s, p, rhs)
    ref = _loop_measure(s, p.astype(np.float64), rhs.astype(np.float64))
    assert abs(got - ref) <= 1e-12 * ref


def _loop_operator(s, p):
    w = s.workload
    c = s.coeffs
    ax, by, cz = c["_cn1_terms"]
    out = np.zeros((w.im + 2, w.jm + 2, w.km + 2))
    for k in range(1, w.km + 1):
        for j in range(1, w.jm + 1):
            for i in range(1, w.im + 1):
                cn = float(np.float32(1.0) / np.float32(np.float32(ax[i - 1] + by[j - 1]) + cz[k - 1]))
                out[i, j, k] = (float(c["cn2l"][i - 1]) * p[i + 1, j, k] + float(c["cn2s"][i - 1]) * p[i - 1, j, k]
                                + float(c["cn3l"][j - 1]) * p[i, j + 1, k] + float(c["cn3s"][j - 1]) * p[i, j - 1, k]
                                + float(c["cn4l"][k - 1]) * p[i, j, k + 1]
                                + float(c["cn4s"][k - 1]) * p[i, j, k - 1] - p[i, j, k] / cn)
    return out


def test_defect_measure_vanishes_on_a_solution():
    """rhs = B(p) gives 0 to FP64 rounding, and 1/cn1 is the sum of the six
    coefficients (init.f95:82-107) to FP32 rounding, so B maps a constant
    to (nearly) 0."""
    s = setup_gen.build(setup_gen.small_workload(im=6, jm=4, km=5, seed=4), dense=False)
    w = s.workload
    rng = np.random.default_rng(1)
    p = rng.uniform(-1, 1, layout.new_field("p", w.im, w.jm, w.km).shape)
    rhs = _loop_operator(s, p)
    b = _bench()
    scale = b.defect_measure(s, p, np.zeros_like(rhs))
    assert scale > 0.0
    assert b.defect_measure(s, p, rhs) <= 1e-24 * scale
    c = s.coeffs
    ax, by, cz = c["_cn1_terms"]
    for k in range(w.km):
        for j in range(w.jm):
            for i in range(w.im):
                tot = (float(c["cn2l"][i]) + float(c["cn2s"][i]) + float(c["cn3l"][j]) + float(c["cn3s"][j])
                       + float(c["cn4l"][k]) + float(c["cn4s"][k]))
                inv = float(np.float32(np.float32(ax[i] + by[j]) + cz[k]))
                assert abs(tot - inv) <= 1e-5 * inv, (i, j, k, tot, inv)
