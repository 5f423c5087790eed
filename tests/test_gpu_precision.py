"""The FP64 defect-correction press (press mode 1) on the GPU.

K inner iterations from e = 0 on the defect equation are, in exact
arithmetic, K further iterations of the same (linear, stationary) SOR, so
outer x inner iterations of the variant are compared with as many plain
iterations:

* on a uniform grid with a right-hand side chosen through u, 10 x 100
  iterations give the least-squares solution of the Poisson matrix (built
  independently; Neumann in i and k, periodic in j) to the FP32 rounding of
  the returned p, while 1000 FP32 SOR iterations stop at a larger error; the
  defect norm falls at every outer step, by 1e-18 overall;
* from the same FP32 inputs, 10 x 100 iterations land closer to the FP64
  restatement's 1000 FP64 iterations (orc64) than 1000 FP32 iterations do.
"""
import dataclasses

import numpy as np
import pytest

from backends import rel_l2
from lesamd import api, layout, setup_gen
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_defect_correction_solves_the_discrete_poisson_problem_to_fp64():
    from test_oracle_math import _poisson_matrix, _uniform_domain
    s = _uniform_domain(8, 6, 5, nmaxp=1000, pjuge=1e-30)
    w = s.workload
    print("\n[defect correction] direct solve 8x6x5", flush=True)
    rng = np.random.default_rng(0)
    r = rng.standard_normal((w.im, w.jm, w.km))
    dx1 = s.metrics["dx1"][2: w.im + 2].astype(np.float64)
    u = layout.new_field("u", w.im, w.jm, w.km)
    col = lambda i: layout.ref_slice("u", (i, 1, 1), (i, w.jm, w.km))  # noqa: E731
    acc = np.zeros((w.jm, w.km))
    for i in range(1, w.im + 1):
        acc = acc + r[i - 1] * w.dt * dx1[i - 1]
        u[col(i)] = acc[None].astype(np.float32)
    uf = u.astype(np.float64)
    rr = np.stack([(uf[col(i)] - uf[col(i - 1)])[0] / dx1[i - 1] / w.dt for i in range(1, w.im + 1)])
    x = np.linalg.lstsq(_poisson_matrix(s), -(rr - rr.mean()).reshape(-1, order="F"), rcond=None)[0]
    x = x - x.mean()
    errs = {}
    for mode in (0, 1):
        d = api.Domain(s)
        if mode:
            d.set_press_mode(1, outer=10, inner=100)
        d.upload("u", u)
        d.stage("press", 1, 1)
        p = layout.interior("p", d.download("p"), w.im, w.jm, w.km).astype(np.float64)
        errs[mode] = np.abs(p.reshape(-1, order="F") - x).max() / np.abs(x).max()
        if mode:
            st = d.press_stats()
            assert st["iters"] == 10
            h = st["sor"]
            assert np.all(np.diff(h[:6]) < 0) and h[-1] <= 1e-18 * h[0], h
        d.close()
    print(f"max error: FP32 SOR {errs[0]:.2e}, defect correction {errs[1]:.2e}", flush=True)
    # the result is rounded to FP32 at the end: a few ulp of the largest value
    assert errs[1] <= 3e-7 and errs[1] < errs[0]


def test_defect_correction_matches_the_fp64_restatement():
    base = setup_gen.small_workload(24, 20, 16, seed=41)
    w = dataclasses.replace(base, nmaxp=1000, pjuge=1e-30)
    s = setup_gen.build(w)
    print("\n[defect correction] vs orc64, 24x20x16", flush=True)
    # FP32 inputs of press: the state after adam of step 2
    src = api.Domain(s)
    src.run(1, 1, 2)
    for st in ("velnw", "bondv1", "velfg", "feedbf", "les", "adam"):
        src.stage(st, 2, 2)
    names = ["u", "v", "w", "f", "g", "h", "p"]
    state = {n: src.download(n) for n in names}
    src.close()
    o64 = oracle.OracleBackend(s, 64, redmode=1)
    o64.load(state)
    o64.call("press", n=2, nmax=2)
    p64 = o64.fetch(["p"])["p"]
    got = {}
    for mode in (0, 1):
        d = api.Domain(s)
        if mode:
            d.set_press_mode(1, outer=10, inner=100)
        for n in names:
            d.upload(n, state[n])
        d.stage("press", 2, 2)
        got[mode] = d.download("p")
        d.close()
    e0, e1 = rel_l2(got[0], p64), rel_l2(got[1], p64)
    print(f"rel L2 vs orc64: FP32 SOR {e0:.2e}, defect correction {e1:.2e}", flush=True)
    assert e1 <= 0.5 * e0


def _env_domain(s, **env):
    import os
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


@pytest.mark.parametrize("face", ["i0", "im1", "j0", "jm1", "k0", "km1"])
def test_defect_correction_ignores_the_ghosts_of_p(face):
    """Press mode 1 reads p only on the interior (its defect mirrors the
    interior in i and k and wraps j), so ghosts of p that differ from their
    boundp copies -- an upload of p, or a restart file without ghosts --
    must not change a bit of the result.  Compared with the same press on
    press-consistent ghosts and with the one-row SOR kernel
    (LES_SOR_ROWS=1), whose iterates equal the two-row kernel's."""
    w = setup_gen.Workload("dc-ghosts", 48, 40, 18, seed=5)
    s = setup_gen.build(w, dense=False)
    print(f"\n[defect correction] perturbed ghost face {face}, 48x40x18", flush=True)
    src = api.Domain(s)
    src.run(1, 3, 8)
    for st in ("velnw", "bondv1", "velfg", "feedbf", "les", "adam"):
        src.stage(st, 4, 8)
    state = {n: src.download(n) for n in ("u", "v", "w", "p", "f", "g", "h")}
    src.close()
    im, jm, km = w.im, w.jm, w.km
    bad = state["p"].copy()  # p(0:im+2, 0:jm+2, 0:km+1)
    rng = np.random.default_rng(3)
    sl = {"i0": np.s_[0, 1:jm + 1, 1:km + 1], "im1": np.s_[im + 1, 1:jm + 1, 1:km + 1],
          "j0": np.s_[1:im + 1, 0, 1:km + 1], "jm1": np.s_[1:im + 1, jm + 1, 1:km + 1],
          "k0": np.s_[1:im + 1, 1:jm + 1, 0], "km1": np.s_[1:im + 1, 1:jm + 1, km + 1]}[face]
    bad[sl] += rng.uniform(0.5, 1.0, bad[sl].shape).astype(np.float32)
    out = {}
    for name, p, env in (("clean", state["p"], {}), ("perturbed", bad, {}),
                         ("rows1", bad, {"LES_SOR_ROWS": 1})):
        d = _env_domain(s, **env)
        d.set_press_mode(1, outer=3, inner=7)
        for n, a in state.items():
            d.upload(n, p if n == "p" else a)
        d.stage("press", 4, 8)
        out[name] = {f: d.download(f) for f in ("p", "rhs", "f", "g", "h")}
        out[name]["hist"] = d.press_stats()["sor"]
        d.close()
    for name in ("perturbed", "rows1"):
        for f in ("p", "rhs", "f", "g", "h"):
            assert np.array_equal(out[name][f], out["clean"][f]), (name, f)
        np.testing.assert_array_equal(out[name]["hist"], out["clean"]["hist"])
