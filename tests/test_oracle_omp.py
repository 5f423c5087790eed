"""The OpenMP build of the restatement (oracle/liblesoracle_omp.so), used
only as the multi-core CPU timing baseline, computes what the serial build
computes: the stencil stages bit for bit (their k loops are split over
threads, each cell written from arrays the loop does not write), press to
the rounding of its FP64 sums, which the threads reduce in another order."""
import os

import numpy as np
import pytest

from backends import rel_l2
from lesamd import setup_gen
from oracle import oracle

FIELDS = ["u", "v", "w", "f", "g", "h", "usum", "sm", "fold"] + [f"diu{q}" for q in range(1, 10)]


def test_openmp_build_equals_serial_build():
    if not oracle.oracle_lib_path(omp=True).exists():
        pytest.skip("oracle/liblesoracle_omp.so not built (make -C oracle)")
    os.environ.setdefault("OMP_NUM_THREADS", "4")
    s = setup_gen.build(setup_gen.small_workload(40, 36, 24, seed=61))
    ser = oracle.OracleBackend(s, 32, redmode=1)
    par = oracle.OracleBackend(s, 32, redmode=1, omp=True)
    for n in (1, 2, 3):
        for st in oracle.STAGES:
            kw = {"n": n, "nmax": 3} if st in ("bondv1", "adam", "press") else {}
            ser.call(st, **kw)
            par.call(st, **kw)
            if st != "press":
                a, b = ser.fetch(FIELDS), par.fetch(FIELDS)
                for f in FIELDS:
                    assert np.array_equal(a[f], b[f]), (n, st, f)
        a, b = ser.fetch(["p", "rhs"]), par.fetch(["p", "rhs"])
        assert rel_l2(b["p"], a["p"]) <= 1e-6 and rel_l2(b["rhs"], a["rhs"]) <= 1e-6
        assert ser.press_stats()["iters"] == par.press_stats()["iters"]
