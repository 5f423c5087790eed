"""Host-side array layouts of the reference's hot path.

Every array of the per-step path is a Fortran explicit-shape array with
non-unit lower bounds, i fastest (main.f95:63-189 and the callee dummies,
e.g. velnw.f95:7-21, vel2.f95:7-47, les.f95:9-33, press.f95:11-39).  Host
code keeps exactly those layouts (numpy, order='F') so arrays can be handed
to the C ABI, which takes a pointer to the lower-bound corner, exactly as a
Fortran caller passes them.  Extents use ip,jp,kp; the reference requires
ip,jp,kp == im,jm,km at run time (ifdata.f95:238-241).
"""
from __future__ import annotations

import numpy as np

# kind -> (lower bounds, extents as a function of (ip, jp, kp))
KINDS = {
    "U": ((0, -1, 0), lambda ip, jp, kp: (ip + 2, jp + 3, kp + 2)),   # u, v, cmask1
    "W": ((0, -1, -1), lambda ip, jp, kp: (ip + 2, jp + 3, kp + 3)),  # w
    "P": ((0, 0, 0), lambda ip, jp, kp: (ip + 3, jp + 3, kp + 2)),    # p
    "F": ((0, 0, 0), lambda ip, jp, kp: (ip + 1, jp + 1, kp + 1)),    # f,g,h,usum..,fx..
    "I": ((1, 1, 1), lambda ip, jp, kp: (ip, jp, kp)),                # fold..,fghold,cn1
    "R": ((0, 0, 0), lambda ip, jp, kp: (ip + 2, jp + 2, kp + 2)),    # rhs, amask1, dmask1
    "S": ((-1, -1, 0), lambda ip, jp, kp: (ip + 3, jp + 3, kp + 2)),  # sm
    "D1": ((-1, 0, 0), lambda ip, jp, kp: (ip + 4, jp + 3, kp + 3)),  # diu1,diu5,cov1,cov5,nou1,nou5
    "D": ((0, 0, 0), lambda ip, jp, kp: (ip + 3, jp + 3, kp + 3)),    # other diu/cov/nou
    "B": ((-1, 0, 0), lambda ip, jp, kp: (ip + 3, jp + 2, kp + 2)),   # bmask1
    "2D": ((0, 0), lambda ip, jp, kp: (ip + 2, jp + 2)),              # uspd, vspd
    "DU": ((0, 1, 1), lambda ip, jp, kp: (ip + 1, jp, kp)),           # dfu1
    "DV": ((1, 0, 1), lambda ip, jp, kp: (ip, jp + 1, kp)),           # dfv1
    "AU": ((1, 1, 0), lambda ip, jp, kp: (ip, jp, kp + 1)),           # aveu, avev (aveflow.f95:18,20)
    "AW": ((1, 1, 0), lambda ip, jp, kp: (ip + 1, jp, kp + 3)),       # avew (aveflow.f95:22)
}

FIELD_KIND = {
    "u": "U", "v": "U", "w": "W", "p": "P",
    "f": "F", "g": "F", "h": "F", "usum": "F", "vsum": "F", "wsum": "F",
    "fx": "F", "fy": "F", "fz": "F",
    "fold": "I", "gold": "I", "hold": "I", "fghold": "I", "cn1": "I", "dfw1": "I",
    "rhs": "R", "amask1": "R", "dmask1": "R", "bmask1": "B", "cmask1": "U",
    "sm": "S", "uspd": "2D", "vspd": "2D", "dfu1": "DU", "dfv1": "DV",
    "aveu": "AU", "avev": "AU", "avew": "AW",
}
# the time statistics of aveflow (aveflow.f95:8-23), in its argument order
AVE_FIELDS = ("aveu", "avev", "avew", "avep", "avel", "aveuu", "avevv", "aveww", "avesm", "avesmsm", "uwfx")
for _n in AVE_FIELDS[3:]:
    FIELD_KIND[_n] = "I"
for _n in range(1, 10):
    _k = "D1" if _n in (1, 5) else "D"
    for _p in ("diu", "cov", "nou"):
        FIELD_KIND[f"{_p}{_n}"] = _k

# 1-D metric arrays: name -> (lower bound, extent as a function of (ip, jp, kp))
METRIC_KIND = {
    "dx1": (-1, lambda ip, jp, kp: ip + 3),   # dx1(-1:ip+1)
    "dxl": (0, lambda ip, jp, kp: ip + 1),    # dxl(0:ip)
    "dxs": (0, lambda ip, jp, kp: ip + 1),    # dxs(0:ip)
    "dy1": (0, lambda ip, jp, kp: jp + 2),    # dy1(0:jp+1)
    "dyl": (0, lambda ip, jp, kp: jp + 1),
    "dys": (0, lambda ip, jp, kp: jp + 1),
    "dzn": (-1, lambda ip, jp, kp: kp + 4),   # dzn(-1:kp+2)
    "dzs": (-1, lambda ip, jp, kp: kp + 4),
    "z2": (0, lambda ip, jp, kp: kp + 3),     # z2(0:kp+2)
    "cn2l": (1, lambda ip, jp, kp: ip), "cn2s": (1, lambda ip, jp, kp: ip),
    "cn3l": (1, lambda ip, jp, kp: jp), "cn3s": (1, lambda ip, jp, kp: jp),
    "cn4l": (1, lambda ip, jp, kp: kp), "cn4s": (1, lambda ip, jp, kp: kp),
    "delx1": (1, lambda ip, jp, kp: kp),
}

# The prognostic and per-step arrays a domain carries between steps.
STEP_FIELDS = (
    "u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold",
    "fghold", "rhs", "sm", "uspd", "vspd", "fx", "fy", "fz", "dfu1", "dfv1", "dfw1",
) + tuple(f"{p}{n}" for p in ("diu", "cov", "nou") for n in range(1, 10))


def field_shape(name: str, ip: int, jp: int, kp: int):
    lb, ext = KINDS[FIELD_KIND[name]]
    return lb, ext(ip, jp, kp)


def new_field(name: str, ip: int, jp: int, kp: int, dtype=np.float32) -> np.ndarray:
    _, shape = field_shape(name, ip, jp, kp)
    return np.zeros(shape, dtype=dtype, order="F")


def new_metric(name: str, ip: int, jp: int, kp: int, dtype=np.float32) -> np.ndarray:
    _, ext = METRIC_KIND[name]
    return np.zeros(ext(ip, jp, kp), dtype=dtype)


def lower_bounds(name: str):
    if name in FIELD_KIND:
        return KINDS[FIELD_KIND[name]][0]
    return (METRIC_KIND[name][0],)


def ref_slice(name: str, lo, hi):
    """numpy slices selecting reference indices lo..hi (inclusive) per axis."""
    lb = lower_bounds(name)
    return tuple(slice(l - b, h - b + 1) for l, h, b in zip(lo, hi, lb))


def at(name: str, arr: np.ndarray, *idx):
    """Element at reference indices idx."""
    lb = lower_bounds(name)
    return arr[tuple(i - b for i, b in zip(idx, lb))]


def interior(name: str, arr: np.ndarray, im: int, jm: int, km: int) -> np.ndarray:
    """The cells 1..im, 1..jm, 1..km of a 3-D field (a view)."""
    return arr[ref_slice(name, (1, 1, 1), (im, jm, km))]


def embed(name: str, small: np.ndarray, big: np.ndarray, im: int, jm: int, km: int) -> None:
    """Copy an array laid out for extents (im,jm,km) into the same reference
    index range of an array laid out for larger compile-time extents."""
    if small.ndim == 1:
        big[: small.shape[0]] = small
        return
    sl = tuple(slice(0, s) for s in small.shape)
    big[sl] = small


def extract(name: str, big: np.ndarray, im: int, jm: int, km: int) -> np.ndarray:
    """Inverse of embed: the (im,jm,km)-extent view of a larger array."""
    if name in FIELD_KIND:
        _, shape = field_shape(name, im, jm, km)
    else:
        shape = (METRIC_KIND[name][1](im, jm, km),)
    sl = tuple(slice(0, s) for s in shape)
    return np.asfortranarray(big[sl])
