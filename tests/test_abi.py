"""The C ABI library loads and exports every function include/les_amd.h
declares; host-side logic that needs no GPU."""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "les_amd.h"
LIB = ROOT / "mpi-les_amd" / "lesamd" / "liblesamd.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(les_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        subprocess.run([sys.executable, str(ROOT / "mpi-les_amd" / "build.py")], check=True)
    return ctypes.CDLL(str(LIB))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for stage in ("velnw", "bondv1", "velfg", "feedbf", "les", "adam", "press"):
        assert f"les_{stage}" in names


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_signatures_match_python_binding():
    """The Python layer passes exactly the reference argument lists."""
    from lesamd import api
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    for stage, names in api.SIGNATURES.items():
        m = re.search(rf"int les_{stage}\((.*?)\);", text, flags=re.S)
        params = [p.strip().split()[-1].lstrip("*") for p in m.group(1).split(",")]
        assert params == names, stage


def test_library_reports_missing_device_cleanly(lib):
    """Without a usable GPU the library returns an error, it does not abort."""
    from lesamd import api, setup_gen
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    s = setup_gen.build(setup_gen.small_workload())
    with pytest.raises(api.LesError):
        api.Domain(s)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_calls_on_unknown_or_destroyed_handles_are_refused(lib):
    """Every domain and batch entry point checks its handle against the live
    ones, so a call on a destroyed (or never created) domain returns an
    error instead of touching freed memory; no GPU is needed to reach it."""
    lib.les_last_error.restype = ctypes.c_char_p
    bogus = ctypes.c_void_p(0x1234560)
    for name, args in (("les_domain_step", (bogus, 2, 10)), ("les_domain_destroy", (bogus,)),
                       ("les_domain_synchronize", (bogus,)), ("les_batch_step", (bogus, None, None)),
                       ("les_batch_destroy", (bogus,)), ("les_batch_query", (bogus,))):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        rc = f(*args)
        assert rc != 0, name
        assert b"destroyed" in lib.les_last_error() or b"unknown" in lib.les_last_error(), name
    lib.les_domain_destroy.restype = ctypes.c_int
    assert lib.les_domain_destroy(None) == 0  # NULL is a no-op, as free(NULL)
