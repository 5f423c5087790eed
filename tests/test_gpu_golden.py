"""GPU path against the fixtures written from the reference's own Fortran
(tests/golden/), against a direct solve of the discrete Poisson problem, and
on the benchmark workloads against the C restatement.

Tolerances: the reference reduces rhsav, sum(reltmp^2) and pav in FP32
sequential order, the GPU in FP64, so fields after press differ from the
reference at rounding level; per-step relative L2 <= 1e-5 (the north-star
bound) on u, v, w, p, sm and equal SOR iteration counts.  Against the C
restatement in its FP64-reduction mode (the GPU's arithmetic) the stencil
stages are bit for bit (tests/test_gpu_parity.py).
"""
import dataclasses
from pathlib import Path

import numpy as np
import pytest

from backends import GpuBackend, rel_l2
from lesamd import api, layout, setup_gen
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
SMALL = {"even": dict(im=16, jm=12, km=10, seed=21), "odd": dict(im=13, jm=11, km=9, seed=22)}


@pytest.mark.parametrize("name", sorted(SMALL))
def test_stage_fixtures(name):
    """Every stage of steps 1-2 against the reference's outputs: bit for bit
    up to press, press within FP32 rounding of the FP64 reductions."""
    fx = np.load(GOLDEN / f"stages_{name}.npz")
    s = setup_gen.build(setup_gen.small_workload(ustar=None, **SMALL[name]))
    print(f"\n[golden stages] {name}", flush=True)
    gpu = GpuBackend(s)
    for n in (1, 2):
        for st in oracle.STAGES:
            kw = {"n": n, "nmax": 1000} if st in ("bondv1", "adam", "press") else {}
            gpu.call(st, **kw)
            keys = [k for k in fx.files if k.startswith(f"{n}/{st}/")]
            got = gpu.fetch([k.split("/")[2] for k in keys])
            for k in keys:
                a, want = got[k.split("/")[2]], fx[k]
                if n == 1 and st != "press":
                    assert np.array_equal(a, want), k
                else:
                    assert rel_l2(a, want) <= 1e-5, (k, rel_l2(a, want))


def test_c0_trajectory_against_reference_norms():
    """150x150x90 plain workload (reference inflow), 20 steps: norms, centre
    values and SOR iteration counts of the reference."""
    fx = np.load(GOLDEN / "c0_norms.npz")
    w = dataclasses.replace(setup_gen.workloads()["plain-150x150x90"], ustar=None)
    s = setup_gen.build(w)
    print(f"\n[golden c0] {w.name} 20 steps", flush=True)
    dom = api.Domain(s)
    dom.set_graph(True)
    for n in range(1, 21):
        dom.step(n, 20)
        st = dom.press_stats()
        assert st["iters"] == fx["iters"][n - 1], n
        # The reference adds 2e6 squares per colour sequentially in FP32:
        # each addition to a large running sum drops low bits, so its sum is
        # biased low by ~6e-4 here.  The GPU's FP64 sum is the accurate one
        # (it matches the FP64-accumulating restatement to 1e-9,
        # tests/test_gpu_parity.py); here only agreement to 5e-3.
        np.testing.assert_allclose(st["sor"], fx["sor"][n - 1][: st["iters"]], rtol=5e-3)
        for k in ("u", "v", "w", "p", "sm"):
            a = dom.download(k).astype(np.float64)
            got = np.array([a.sum(), np.sqrt((a * a).sum()), np.abs(a).max()])
            want = fx[f"norm_{k}"][n - 1]
            # the L2 norm and max are scale-free; the sum of p is ~0 by construction
            assert abs(got[1] - want[1]) <= 1e-5 * want[1], (n, k, got, want)
            assert abs(got[2] - want[2]) <= 1e-5 * want[2], (n, k, got, want)
        c = (w.im // 2, w.jm // 2, w.km // 2)
        cen = [layout.at(k, dom.download(k), *c) for k in ("u", "v", "w")]
        np.testing.assert_allclose(cen, fx["centre"][n - 1][1:], rtol=1e-4, atol=1e-6)
    dom.close()


def test_sor_on_gpu_solves_the_discrete_poisson_problem():
    """Uniform grid, rhs chosen through u: after 1000 iterations the GPU
    pressure equals the least-squares solution of the Poisson matrix built
    independently (Neumann in i,k; periodic in j)."""
    from test_oracle_math import _poisson_matrix, _uniform_domain
    s = _uniform_domain(8, 6, 5, nmaxp=1000, pjuge=1e-30)
    w = s.workload
    print("\n[gpu sor direct solve] 8x6x5", flush=True)
    gpu = GpuBackend(s)
    rng = np.random.default_rng(0)
    r = rng.standard_normal((w.im, w.jm, w.km))
    u = gpu.arrays["u"]
    dx1 = s.metrics["dx1"][2: w.im + 2].astype(np.float64)
    col = lambda i: layout.ref_slice("u", (i, 1, 1), (i, w.jm, w.km))  # noqa: E731
    acc = np.zeros((w.jm, w.km))
    for i in range(1, w.im + 1):
        acc = acc + r[i - 1] * w.dt * dx1[i - 1]
        u[col(i)] = acc[None].astype(np.float32)
    gpu.call("press", n=1, nmax=1)
    st = gpu.press_stats()
    # FP32 iterates stop changing before 1000 iterations: the residual is
    # then exactly 0, below the stop threshold
    assert st["iters"] == 1000 or st["sor"][-1] < 1e-30
    p = layout.interior("p", gpu.fetch(["p"])["p"], w.im, w.jm, w.km).astype(np.float64)
    # the rhs the GPU saw is div(u)/dt of the FP32 u
    uf = gpu.arrays["u"].astype(np.float64)
    rr = np.stack([(uf[col(i)] - uf[col(i - 1)])[0] / dx1[i - 1] / w.dt for i in range(1, w.im + 1)])
    rhs = rr - rr.mean()
    x = np.linalg.lstsq(_poisson_matrix(s), -rhs.reshape(-1, order="F"), rcond=None)[0]
    x = x - x.mean()
    assert np.abs(p.reshape(-1, order="F") - x).max() <= 2e-5 * np.abs(x).max()


def test_plain_150_workload_against_restatement():
    """The 150x150x90 workload (log-law inflow, seeded city) for 3 steps
    against the C restatement with FP64 reductions."""
    wname = "plain-150x150x90"
    w = dataclasses.replace(setup_gen.workloads()[wname], nsteps=3)
    s = setup_gen.build(w)
    print(f"\n[workload] {wname}", flush=True)
    dom = api.Domain(s)
    orc = oracle.OracleBackend(s, 32, redmode=1)
    for n in range(1, 4):
        dom.step(n, 3)
        orc.step(n, 3)
        ref = orc.fetch(["u", "v", "w", "p", "sm"])
        for k in ref:
            e = rel_l2(dom.download(k), ref[k])
            assert e <= 1e-5, (n, k, e)
        assert dom.press_stats()["iters"] == orc.press_stats()["iters"]
    dom.close()


def test_plain_300_workload_properties():
    """The benchmarked 300x300x90 workload, 20 steps through graph replay:
    properties that hold at any size -- the inflow planes carry the log-law
    profile exactly, j ghosts are periodic images, p has zero
    volume-weighted mean, every SOR iteration ran (50), the residual is
    finite, and a second domain run eagerly is bit-identical."""
    w = setup_gen.workloads()["plain-300x300x90"]
    s = setup_gen.build(w)
    print(f"\n[workload properties] {w.name}", flush=True)
    a, b = api.Domain(s), api.Domain(s)
    a.set_graph(True)
    for n in range(1, 21):
        a.step(n, 20)
        b.step(n, 20)
    st = a.press_stats()
    assert st["iters"] == 50 and st["nonfinite"] == 0 and np.all(np.isfinite(st["sor"]))
    u = a.download("u")
    prof = s.inflow
    for i in (0, 1):
        np.testing.assert_array_equal(u[layout.ref_slice("u", (i, 1, 1), (i, w.jm, w.km))][0], prof)
    np.testing.assert_array_equal(u[layout.ref_slice("u", (0, 0, 1), (w.im + 1, 0, w.km))][:, 0],
                                  u[layout.ref_slice("u", (0, w.jm, 1), (w.im + 1, w.jm, w.km))][:, 0])
    p = layout.interior("p", a.download("p"), w.im, w.jm, w.km).astype(np.float64)
    dv = 16.0 * s.metrics["dzn"][2: w.km + 2].astype(np.float64)[None, None, :]
    assert abs((p * dv).sum()) <= 1e-6 * (np.abs(p) * dv).sum()
    for k in ("u", "v", "w", "p", "sm", "f", "usum"):
        assert np.array_equal(a.download(k), b.download(k)), k
    a.close()
    b.close()
