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


FORTRAN = ROOT / "mpi-les_amd" / "fortran" / "les_amd_modules.F90"


def _fortran_interfaces():
    """name -> list of C types of the bind(C) interfaces in the drop-in
    modules, from their Fortran declarations: value integer(c_int) -> int,
    value real(c_float) -> float, an assumed-size real(c_float) array ->
    float*, a non-value scalar integer(c_int) / real(c_double) /
    real(c_float) -> int* / double* / float*, type(c_ptr) value -> void*."""
    text = FORTRAN.read_text()
    text = re.sub(r"&\s*\n\s*", " ", text)  # join continuation lines
    out = {}
    for m in re.finditer(r"function\s+(\w+)\s*\(([^)]*)\)\s*bind\(C,\s*name='(\w+)'\)(.*?)end function", text,
                         flags=re.S | re.I):
        args = [a.strip() for a in m.group(2).split(",") if a.strip()]
        body = m.group(4)
        kinds = {}
        for dm in re.finditer(r"^\s*(integer\(c_int\)|real\(c_float\)|real\(c_double\)|type\(c_ptr\))"
                              r"(\s*,\s*value)?\s*::\s*(.*)$", body, flags=re.M | re.I):
            base = {"integer(c_int)": "int", "real(c_float)": "float", "real(c_double)": "double",
                    "type(c_ptr)": "void*"}[dm.group(1).lower()]
            for var in re.findall(r"(\w+)(\(\*\))?", dm.group(3)):
                name, arr = var
                if dm.group(2) or base == "void*":
                    kinds[name.lower()] = base
                else:
                    kinds[name.lower()] = base + "*"
        out[m.group(3)] = [kinds.get(a.lower(), "?" + a) for a in args]
    return out


def _c_prototypes():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for m in re.finditer(r"\bint\s+(les_\w+)\s*\(([^)]*)\);", text, flags=re.S):
        types = []
        for p in m.group(2).split(","):
            p = p.strip().replace("const ", "")
            if not p or p == "void":
                continue
            t = p.rsplit(None, 1)[0] + ("*" * p.split()[-1].count("*"))
            types.append(t.replace(" ", ""))
        out[m.group(1)] = types
    return out


def test_fortran_interfaces_match_the_c_prototypes():
    """Every bind(C) interface the drop-in Fortran modules declare has the
    argument count and types of the prototype in include/les_amd.h (checked
    from the two sources, without a GPU)."""
    f, c = _fortran_interfaces(), _c_prototypes()
    assert len(f) >= 12
    for name, ftypes in f.items():
        if name == "les_last_error":
            continue
        assert name in c, name
        ctypes_ = c[name]
        assert len(ftypes) == len(ctypes_), (name, ftypes, ctypes_)
        for ft, ct in zip(ftypes, ctypes_):
            # a C void* parameter takes any array (les_sync_host's host array)
            assert ft == ct or (ct == "void*" and ft.endswith("*")), (name, ftypes, ctypes_)


def test_fortran_modules_compile_with_verbose(tmp_path):
    """The VERBOSE prints of the press module (press.f95:164-195) compile."""
    flang = Path("/opt/rocm/lib/llvm/bin/flang")
    if not flang.exists():
        pytest.skip("flang absent")
    r = subprocess.run([str(flang), "-cpp", "-O0", "-ffree-form", "-DVERBOSE", "-module-dir", str(tmp_path),
                        "-c", str(FORTRAN), "-o", str(tmp_path / "m.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
