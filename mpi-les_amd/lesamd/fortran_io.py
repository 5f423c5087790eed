"""Fortran unformatted sequential files as the reference writes them: each
record is a 4-byte length, the payload, and the length again, in the byte
order of the build (big-endian in the reference's build, SConstruct:77).

Readers for the dumps of the hot path:
* data20 (press.f95:205-217): u, v, w, p, usum, vsum, wsum over the interior
  1..im x 1..jm x 1..km, i fastest, one record each;
* data21 (adam.f95:20-29): fold, gold, hold, fghold, likewise;
* data10 and data11 (aveflow.f95:207-345): the time statistics aveu, avew,
  avev and aveuu, aveww, avevv, uwfx over the interior, likewise.

And the inflow records bondv1_data24 reads each step
(bondv1_data24.f95:43-68): data24_out<n+48000>.dat, unformatted direct
access, no record markers, 1200 values per record: km records of u(j,k),
then km of w, then km of v; the inflow plane is values 400+j, j=1..jm.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

DATA20 = ("u", "v", "w", "p", "usum", "vsum", "wsum")
DATA21 = ("fold", "gold", "hold", "fghold")
DATA10 = ("aveu", "avew", "avev")
DATA11 = ("aveuu", "aveww", "avevv", "uwfx")
DATA24_RECORD = 1200  # values per record (recl=4*1200, bondv1_data24.f95:45)
DATA24_OFFSET = 400   # plane j is record value 400+j (bondv1_data24.f95:64-66)


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


def read_data10(path, im, jm, km, byteorder: str = ">") -> dict:
    return _fields(path, DATA10, im, jm, km, byteorder)


def read_data11(path, im, jm, km, byteorder: str = ">") -> dict:
    return _fields(path, DATA11, im, jm, km, byteorder)


def ave_interior(name: str, a: np.ndarray, km: int) -> np.ndarray:
    """The (im,jm,km) interior aveflow writes from a statistics array in
    its reference layout: aveu/avev (ip,jp,0:kp), avew (ip+1,jp,0:kp+2),
    the rest (ip,jp,kp)."""
    if name in ("aveu", "avev"):
        return a[:, :, 1:km + 1]
    if name == "avew":
        return a[:-1, :, 1:km + 1]
    return a


def write_aveflow(path10, path11, stats: dict, km: int, byteorder: str = ">") -> None:
    """data10 and data11 as aveflow.f95:207-345 writes them, from the
    statistics in reference layout (Domain.download("aveu") etc.)."""
    inner = {n: ave_interior(n, stats[n], km) for n in DATA10 + DATA11}
    write_fields(path10, inner, DATA10, byteorder)
    write_fields(path11, inner, DATA11, byteorder)


def data24_name(n: int) -> str:
    """File name of step n's inflow records (bondv1_data24.f95:43)."""
    return f"data24_out{n + 48000:06d}.dat"


def write_data24(path, ub: np.ndarray, vb: np.ndarray, wb: np.ndarray, byteorder: str = ">") -> None:
    """Records of one step from full-width planes (1200, km): u, then w, then v."""
    dt = np.dtype(f"{byteorder}f4")
    km = ub.shape[1]
    with open(path, "wb") as fh:
        for a in (ub, wb, vb):
            if a.shape != (DATA24_RECORD, km):
                raise ValueError(f"data24 plane of shape {a.shape}, expected {(DATA24_RECORD, km)}")
            fh.write(np.asarray(a, np.float32).astype(dt).tobytes(order="F"))


def read_data24(path, jm: int, km: int, byteorder: str = ">"):
    """The inflow planes (u, v, w), each (jm, km), of one data24 file."""
    if DATA24_OFFSET + jm > DATA24_RECORD:
        raise ValueError(f"jm = {jm} does not fit the {DATA24_RECORD}-value records")
    raw = np.fromfile(path, dtype=np.dtype(f"{byteorder}f4"))
    if raw.size < 3 * km * DATA24_RECORD:
        raise ValueError(f"{path}: {raw.size} values, expected at least {3 * km * DATA24_RECORD}")
    rec = raw[: 3 * km * DATA24_RECORD].reshape((DATA24_RECORD, 3 * km), order="F").astype(np.float32)
    sl = slice(DATA24_OFFSET, DATA24_OFFSET + jm)
    ua = np.asfortranarray(rec[sl, 0:km])
    wa = np.asfortranarray(rec[sl, km:2 * km])
    va = np.asfortranarray(rec[sl, 2 * km:3 * km])
    return ua, va, wa
