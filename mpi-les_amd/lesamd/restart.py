"""Restart files and checkpoints of a device-resident domain.

* The reference's restart pair (anime.f95:326-595 writes it at n = n1-1 or
  n = nmax; ifdata.f95:252-469 reads it when ical == 1):
    data30NNNNNN.dat: record (n: int32, time: real32), then u, v, w, p, usum,
                      vsum, wsum over the interior 1..im x 1..jm x 1..km
    data31NNNNNN.dat: f, g, h, fold, gold, hold
  unformatted sequential, big-endian (SConstruct:77).  Reading one back sets
  the interior cells only, as the ical == 1 path does; ghost cells keep
  their values.
* A checkpoint: every field the step reads or carries to the next step,
  ghost cells included, so a domain restored from it continues bit for bit
  as the saved one would have.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import fortran_io, layout

DATA30 = ("u", "v", "w", "p", "usum", "vsum", "wsum")
DATA31 = ("f", "g", "h", "fold", "gold", "hold")
# the state a step carries to the next (everything else is rebuilt each step)
CHECKPOINT_FIELDS = ("u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "rhs",
                     "sm", "uspd", "vspd") + tuple(f"diu{q}" for q in range(1, 10))


def restart_names(n: int) -> tuple[str, str]:
    """File names the reference gives the pair at step n (anime.f95:375, 466)."""
    return f"data30{n:06d}.dat", f"data31{n:06d}.dat"


def write_restart(directory, n: int, time: float, fields: dict, im: int, jm: int, km: int,
                  byteorder: str = ">") -> tuple[Path, Path]:
    """fields: arrays in reference layout (ghosts included); the interiors
    are written."""
    d = Path(directory)
    p30, p31 = (d / x for x in restart_names(n))
    head = np.array([n], f"{byteorder}i4").tobytes() + np.array([time], f"{byteorder}f4").tobytes()
    dt = np.dtype(f"{byteorder}f4")

    def rec(name):
        return np.asarray(layout.interior(name, fields[name], im, jm, km), np.float32).astype(dt).tobytes(order="F")

    fortran_io.write_records(p30, [head] + [rec(nm) for nm in DATA30], byteorder)
    fortran_io.write_records(p31, [rec(nm) for nm in DATA31], byteorder)
    return p30, p31


def read_restart(path30, path31, im: int, jm: int, km: int, byteorder: str = ">"):
    """(n, time, {name: interior (im, jm, km)}) from a restart pair."""
    r30 = fortran_io.read_records(path30, byteorder)
    if len(r30) != 1 + len(DATA30) or len(r30[0]) != 8:
        raise ValueError(f"{path30}: not a data30 restart file")
    n = int(np.frombuffer(r30[0], f"{byteorder}i4", 1, 0)[0])
    time = float(np.frombuffer(r30[0], f"{byteorder}f4", 1, 4)[0])
    out = {}
    for name, r in zip(DATA30, r30[1:]):
        out[name] = _field(path30, name, r, im, jm, km, byteorder)
    r31 = fortran_io.read_records(path31, byteorder)
    if len(r31) != len(DATA31):
        raise ValueError(f"{path31}: not a data31 restart file")
    for name, r in zip(DATA31, r31):
        out[name] = _field(path31, name, r, im, jm, km, byteorder)
    return n, time, out


def _field(path, name, rec, im, jm, km, byteorder):
    a = np.frombuffer(rec, np.dtype(f"{byteorder}f4"))
    if a.size != im * jm * km:
        raise ValueError(f"{path}: record {name} holds {a.size} values, expected {im * jm * km}")
    return a.astype(np.float32).reshape((im, jm, km), order="F")


def save_restart(dom, directory, n: int, time: float) -> tuple[Path, Path]:
    """Write the restart pair of a domain (its state after step n)."""
    w = dom.workload
    fields = {nm: dom.download(nm) for nm in DATA30 + DATA31}
    return write_restart(directory, n, time, fields, w.im, w.jm, w.km)


def load_restart(dom, path30, path31) -> tuple[int, float]:
    """Set the interiors of a domain from a restart pair (ghost cells keep
    their values, as on the reference's ical == 1 start); returns (n, time)
    of the pair.  The next step to run is n + 1."""
    w = dom.workload
    n, time, fields = read_restart(path30, path31, w.im, w.jm, w.km)
    for name, a in fields.items():
        full = dom.download(name)
        full[layout.ref_slice(name, (1, 1, 1), (w.im, w.jm, w.km))] = a
        dom.upload(name, full)
    return n, time


def save_checkpoint(dom, path, n: int) -> None:
    """Every carried field with its ghost cells, and the step it follows."""
    np.savez(path, n=np.array(n), **{nm: dom.download(nm) for nm in CHECKPOINT_FIELDS})


def load_checkpoint(dom, path) -> int:
    """Restore a checkpoint into a domain of the same set-up; returns the
    step it follows (continue with n + 1)."""
    with np.load(path, allow_pickle=False) as z:
        for nm in CHECKPOINT_FIELDS:
            dom.upload(nm, z[nm])
        return int(z["n"])
