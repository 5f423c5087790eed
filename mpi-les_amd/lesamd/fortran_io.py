"""Fortran unformatted sequential files as the reference writes them: each
record is a 4-byte length, the payload, and the length again, in the byte
order of the build (big-endian in the reference's build, SConstruct:77).

Readers for the dumps of the hot path:
* data20 (press.f95:205-217): u, v, w, p, usum, vsum, wsum over the interior
  1..im x 1..jm x 1..km, i fastest, one record each;
* data21 (adam.f95:20-29): fold, gold, hold, fghold, likewise.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

DATA20 = ("u", "v", "w", "p", "usum", "vsum", "wsum")
DATA21 = ("fold", "gold", "hold", "fghold")


def read_records(path, byteorder: str = ">") -> list[bytes]:
    raw = Path(path).read_bytes()
    out, pos = [], 0
    mark = np.dtype(f"{byteorder}i4")
    while pos < len(raw):
        if pos + 4 > len(raw):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        n = int(np.frombuffer(raw, mark, 1, pos)[0])
        end = pos + 4 + n
        if n < 0 or end + 4 > len(raw) or int(np.frombuffer(raw, mark, 1, end)[0]) != n:
            raise ValueError(f"{path}: bad record at byte {pos} (length {n})")
        out.append(raw[pos + 4:end])
        pos = end + 4
    return out


def write_records(path, records, byteorder: str = ">") -> None:
    mark = np.dtype(f"{byteorder}i4")
    with open(path, "wb") as fh:
        for r in records:
            r = bytes(r)
            m = np.array([len(r)], mark).tobytes()
            fh.write(m + r + m)


def _fields(path, names, im, jm, km, byteorder):
    recs = read_records(path, byteorder)
    if len(recs) != len(names):
        raise ValueError(f"{path}: {len(recs)} records, expected {len(names)}")
    out = {}
    for name, r in zip(names, recs):
        a = np.frombuffer(r, np.dtype(f"{byteorder}f4"))
        if a.size != im * jm * km:
            raise ValueError(f"{path}: record {name} holds {a.size} values, expected {im * jm * km}")
        out[name] = a.astype(np.float32).reshape((im, jm, km), order="F")
    return out


def read_data20(path, im, jm, km, byteorder: str = ">") -> dict:
    return _fields(path, DATA20, im, jm, km, byteorder)


def read_data21(path, im, jm, km, byteorder: str = ">") -> dict:
    return _fields(path, DATA21, im, jm, km, byteorder)


def write_fields(path, fields: dict, names, byteorder: str = ">") -> None:
    """Interior arrays (im,jm,km) as one record each, i fastest."""
    dt = np.dtype(f"{byteorder}f4")
    write_records(path, [np.asarray(fields[n], np.float32).astype(dt).tobytes(order="F") for n in names],
                  byteorder)
