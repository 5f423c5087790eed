"""The C restatement against the reference's own Fortran subroutines (flang
builds under oracle/_ref), bit for bit, and against the committed golden
fixtures written from those subroutines (tests/golden/make_golden.py).

These tests pin the oracle: every function of les_oracle.c is executed on
seeded inputs and compared with the reference after every stage, in FP32
(against the reference as shipped) and in FP64 (against the kind-promoted
build, the accuracy oracle).
"""
import dataclasses

import numpy as np
import pytest

from lesamd import layout, setup_gen
from oracle import oracle

STEP_ARRAYS = ["u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "rhs",
               "sm", "uspd", "vspd", "delx1"] + [f"diu{q}" for q in range(1, 10)]

REF_CASES = [
    dict(im=24, jm=20, km=16, seed=1),
    dict(im=13, jm=11, km=9, seed=2),                         # odd extents
    dict(im=25, jm=25, km=105, seed=3, dz_rule="reference"),  # the build's full extents, grid.f95 dz rule
    dict(im=9, jm=22, km=30, seed=4, city=None),              # no buildings
]


def _setup(case):
    return setup_gen.build(setup_gen.small_workload(ustar=None, **case))


@pytest.mark.parametrize("case", REF_CASES, ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}")
def test_restatement_equals_reference_every_stage(case, ref_small):
    s = _setup(case)
    ref = oracle.RefBackend(s, ref_small)
    orc = oracle.OracleBackend(s, 32, redmode=0)
    for n in range(1, 4):
        for st in oracle.STAGES:
            kw = {"n": n, "nmax": 10 ** 6} if st in ("bondv1", "adam", "press") else {}
            ref.call(st, **kw)
            orc.call(st, **kw)
            a, b = ref.fetch(STEP_ARRAYS), orc.fetch(STEP_ARRAYS)
            for name in STEP_ARRAYS:
                assert np.array_equal(a[name], b[name]), (n, st, name)


def test_ghost_routines_equal_reference(ref_small):
    """boundp1, boundp2, bondfg, boundsm on random arrays (every ghost)."""
    s = _setup(dict(im=11, jm=9, km=7, seed=5))
    ref = oracle.RefBackend(s, ref_small)
    orc = oracle.OracleBackend(s, 32, redmode=0)
    rng = np.random.default_rng(5)
    state = {n: rng.standard_normal(a.shape).astype(np.float32) for n, a in orc.fetch(["p", "f", "g", "h", "sm"]).items()}
    for fn, names in (("boundp1", ["p"]), ("boundp2", ["p"]), ("bondfg", ["f", "g", "h"]), ("boundsm", ["sm"])):
        ref.load(state)
        orc.load(state)
        ref.call(fn)
        orc.call(fn)
        a, b = ref.fetch(names), orc.fetch(names)
        for n in names:
            assert np.array_equal(a[n], b[n]), (fn, n)


def test_fp64_reduction_mode_equals_mixed_reference_press():
    """press of the reference built with -fdefault-real-8 (FP32 fields, FP64
    scalars and reductions) equals the restatement's FP64-accumulation mode
    (the arithmetic of the GPU path) on the same inputs."""
    if not oracle.ref_available("small_r8"):
        pytest.skip("reference build oracle/_ref/small_r8 absent")
    s = _setup(dict(im=20, jm=16, km=12, seed=6))
    orc = oracle.OracleBackend(s, 32, redmode=1)
    for n in (1, 2):
        orc.step(n, 10 ** 6, stop_after="adam")
        state = orc.fetch(STEP_ARRAYS)
        ref = oracle.RefBackend(s, "small_r8")
        ref.load(state)
        ref.call("press", n=n, nmax=10 ** 6)
        orc.call("press", n=n, nmax=10 ** 6)
        a, b = ref.fetch(["p", "rhs", "f", "g", "h"]), orc.fetch(["p", "rhs", "f", "g", "h"])
        for name in a:
            assert np.array_equal(a[name], b[name]), (n, name)


@pytest.mark.parametrize("case", REF_CASES, ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}")
def test_fp64_restatement_equals_kind_promoted_reference(case):
    """The FP64 instantiation of the restatement (orc64, the accuracy
    oracle) against the reference with every real promoted to real(kind=8)
    (oracle/Makefile, o64_small): bit for bit after every stage of steps
    1-3."""
    if not oracle.ref_available("o64_small"):
        pytest.skip("reference build oracle/_ref/o64_small absent")
    s = _setup(case)
    ref = oracle.RefBackend(s, "o64_small")
    orc = oracle.OracleBackend(s, 64, redmode=0)
    for n in range(1, 4):
        for st in oracle.STAGES:
            kw = {"n": n, "nmax": 10 ** 6} if st in ("bondv1", "adam", "press") else {}
            ref.call(st, **kw)
            orc.call(st, **kw)
            a, b = ref.fetch(STEP_ARRAYS), orc.fetch(STEP_ARRAYS)
            for name in STEP_ARRAYS:
                assert a[name].dtype == np.float64
                assert np.array_equal(a[name], b[name]), (n, st, name)


# --------------------------------------------------- committed fixtures
from pathlib import Path  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"
SMALL = {"even": dict(im=16, jm=12, km=10, seed=21), "odd": dict(im=13, jm=11, km=9, seed=22)}


@pytest.mark.parametrize("name", sorted(SMALL))
def test_restatement_reproduces_stage_fixtures(name):
    fx = np.load(GOLDEN / f"stages_{name}.npz")
    s = _setup(SMALL[name])
    orc = oracle.OracleBackend(s, 32, redmode=0)
    for n in (1, 2):
        for st in oracle.STAGES:
            kw = {"n": n, "nmax": 1000} if st in ("bondv1", "adam", "press") else {}
            orc.call(st, **kw)
            keys = [k for k in fx.files if k.startswith(f"{n}/{st}/")]
            assert keys
            got = orc.fetch([k.split("/")[2] for k in keys])
            for k in keys:
                assert np.array_equal(got[k.split("/")[2]], fx[k]), k


def test_restatement_reproduces_c0_norms_first_steps():
    """150x150x90 plain workload, reference inflow: the first two steps'
    norms, centre values and SOR history (the remaining 18 steps are checked
    on the GPU path)."""
    fx = np.load(GOLDEN / "c0_norms.npz")
    w = dataclasses.replace(setup_gen.workloads()["plain-150x150x90"], ustar=None)
    s = setup_gen.build(w)
    orc = oracle.OracleBackend(s, 32, redmode=0)
    for n in (1, 2):
        orc.step(n, 10 ** 9)
        f = orc.fetch(["u", "v", "w", "p", "sm"])
        for k in ("u", "v", "w", "p", "sm"):
            a = f[k].astype(np.float64)
            np.testing.assert_array_equal([a.sum(), np.sqrt((a * a).sum()), np.abs(a).max()], fx[f"norm_{k}"][n - 1])
        c = (w.im // 2, w.jm // 2, w.km // 2)
        np.testing.assert_array_equal([layout.at(k, f[k], *c) for k in ("p", "u", "v", "w")], fx["centre"][n - 1])
        st = orc.press_stats()
        assert st["iters"] == fx["iters"][n - 1]
        np.testing.assert_array_equal(st["sor"], fx["sor"][n - 1][: st["iters"]])


def test_restatement_reproduces_c1_norms_first_step():
    """300x300x90 plain workload, reference inflow: step 1's norms, centre
    values and SOR history (steps 2-5 are checked on the GPU path; the
    generator compared all 5 steps bit for bit with the reference)."""
    fx = np.load(GOLDEN / "c1_norms.npz")
    w = dataclasses.replace(setup_gen.workloads()["plain-300x300x90"], ustar=None)
    s = setup_gen.build(w)
    orc = oracle.OracleBackend(s, 32, redmode=0)
    orc.step(1, 10 ** 9)
    f = orc.fetch(["u", "v", "w", "p", "sm"])
    for k in ("u", "v", "w", "p", "sm"):
        a = f[k].astype(np.float64)
        np.testing.assert_array_equal([a.sum(), np.sqrt((a * a).sum()), np.abs(a).max()], fx[f"norm_{k}"][0])
    c = (w.im // 2, w.jm // 2, w.km // 2)
    np.testing.assert_array_equal([layout.at(k, f[k], *c) for k in ("p", "u", "v", "w")], fx["centre"][0])
    st = orc.press_stats()
    assert st["iters"] == fx["iters"][0]
    np.testing.assert_array_equal(st["sor"], fx["sor"][0][: st["iters"]])
