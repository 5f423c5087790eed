"""ctypes access to tests/fortran/liblestimeloop.so: the time loop of
main.f95:247-322 compiled against the drop-in Fortran modules."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from lesamd import api, layout

LIB = Path(__file__).resolve().parent / "fortran" / "liblestimeloop.so"
ARRAYS = ("dx1 dy1 dzn dzs dxs dys z2 cn1 cn2l cn2s cn3l cn3s cn4l cn4s bmask1 cmask1 dmask1 "
          "p u v w f g h usum vsum wsum fold gold hold fghold rhs sm uspd vspd delx1").split()


def load():
    api.lib()  # the HIP library first, so a missing build fails with its own message
    L = ctypes.CDLL(str(LIB))
    f = L.les_test_time_loop
    f.restype = None
    f.argtypes = ([ctypes.c_int] * 6 + [ctypes.c_float] * 5 + [ctypes.c_void_p] * len(ARRAYS) +
                  [ctypes.c_char_p] * 2 + [ctypes.c_int, ctypes.c_char_p, ctypes.c_int] +
                  [ctypes.c_void_p] * len(layout.AVE_FIELDS) + [ctypes.c_char_p] * 2)
    return f


def _name(s: str, n: int) -> bytes:
    b = s.encode()
    if len(b) > n:
        raise ValueError(f"{s!r} does not fit character(len={n})")
    return b.ljust(n)


def run(arrays: dict, scalars: dict, n0: int, n1: int, nmax: int, data20: str, data21: str,
        data24_dir: str | None = None, nave1: int = 0, data10: str = "data10", data11: str = "data11") -> dict:
    """Steps n0..n1.  data24_dir: take the inflow from data24 records there
    (bondv1_data24); nave1 > 0: run aveflow after each step from step nave1
    on.  Returns the aveflow statistics arrays (reference layout)."""
    f = load()
    for n in ARRAYS:
        a = arrays[n]
        assert a.dtype == np.float32 and (a.flags.f_contiguous or a.ndim == 1), n
    im, jm, km = scalars["im"], scalars["jm"], scalars["km"]
    ave = {n: layout.new_field(n, im, jm, km) for n in layout.AVE_FIELDS}
    f(im, jm, km, n0, n1, nmax,
      scalars["ro"], scalars["dt"], scalars["vn"], scalars["alpha"], scalars["beta"],
      *[arrays[n].ctypes.data for n in ARRAYS], _name(data20, 70), _name(data21, 70),
      0 if data24_dir is None else 1, _name(data24_dir or "", 512), nave1,
      *[ave[n].ctypes.data for n in layout.AVE_FIELDS], _name(data10, 70), _name(data11, 70))
    return ave
