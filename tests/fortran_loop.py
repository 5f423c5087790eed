"""ctypes access to tests/fortran/liblestimeloop.so: the time loop of
main.f95:247-296 compiled against the drop-in Fortran modules."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from lesamd import api

LIB = Path(__file__).resolve().parent / "fortran" / "liblestimeloop.so"
ARRAYS = ("dx1 dy1 dzn dzs dxs dys z2 cn1 cn2l cn2s cn3l cn3s cn4l cn4s bmask1 cmask1 dmask1 "
          "p u v w f g h usum vsum wsum fold gold hold fghold rhs sm uspd vspd delx1").split()


def load():
    api.lib()  # the HIP library first, so a missing build fails with its own message
    L = ctypes.CDLL(str(LIB))
    f = L.les_test_time_loop
    f.restype = None
    f.argtypes = [ctypes.c_int] * 6 + [ctypes.c_float] * 5 + [ctypes.c_void_p] * len(ARRAYS) + [ctypes.c_char_p] * 2
    return f


def run(arrays: dict, scalars: dict, n0: int, n1: int, nmax: int, data20: str, data21: str):
    f = load()
    for n in ARRAYS:
        a = arrays[n]
        assert a.dtype == np.float32 and (a.flags.f_contiguous or a.ndim == 1), n
    names = [s.encode().ljust(70) for s in (data20, data21)]
    if any(len(s) > 70 for s in names):
        raise ValueError("file names must fit character(len=70)")
    f(scalars["im"], scalars["jm"], scalars["km"], n0, n1, nmax,
      scalars["ro"], scalars["dt"], scalars["vn"], scalars["alpha"], scalars["beta"],
      *[arrays[n].ctypes.data for n in ARRAYS], *names)
