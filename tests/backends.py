"""Test-side driver of the HIP library through its reference-signature C ABI,
with the same interface as the CPU checkers in oracle/oracle.py, so a test
runs identical stage sequences on each and compares arrays."""
from __future__ import annotations

import numpy as np

from lesamd import api, layout
from lesamd.setup_gen import VN


class GpuBackend:
    """Arrays of one domain on the host in reference layout, handed to the
    les_* entry points (which mirror them on the device)."""

    def __init__(self, setup):
        w = setup.workload
        self.w = w
        self.setup = setup
        self.arrays = {n: layout.new_field(n, w.im, w.jm, w.km) for n in layout.STEP_FIELDS}
        for n, a in setup.masks.items():
            self.arrays[n] = np.asfortranarray(a.astype(np.float32))
        for n, a in setup.metrics.items():
            self.arrays[n] = np.ascontiguousarray(a.astype(np.float32))
        for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s"):
            self.arrays[n] = np.ascontiguousarray(setup.coeffs[n].astype(np.float32))
        self.arrays["cn1"] = np.asfortranarray(setup.coeffs["cn1"].astype(np.float32))
        self.arrays["delx1"] = layout.new_metric("delx1", w.im, w.jm, w.km)
        self.scalars = dict(km=w.km, jm=w.jm, im=w.im, ro=1.1763, dt=w.dt, alpha=-10.0, beta=-1.0, vn=float(VN))
        api.release_all()
        api.press_config(w.nmaxp, w.pjuge, w.omega)
        self._prof = api.set_inflow(w.im, w.jm, w.km, setup.inflow)

    def load(self, state: dict):
        for n, a in state.items():
            self.arrays[n][...] = a
            api.mark_host_dirty(self.arrays[n])

    def call(self, stage, **extra):
        args = {}
        for a in api.SIGNATURES[stage]:
            if a in extra:
                args[a] = extra[a]
            elif a in self.scalars:
                args[a] = self.scalars[a]
            else:
                args[a] = self.arrays[a]
        api.call(stage, **args)

    def step(self, n, nmax, stop_after=None):
        for s in api.STAGES:
            kw = {"n": n, "nmax": nmax} if s in ("bondv1", "adam", "press") else {}
            self.call(s, **kw)
            if s == stop_after:
                return

    def fetch(self, names):
        out = {}
        for n in names:
            if n in ("delx1",) or n.startswith(("cov", "nou", "df")) or n in ("fx", "fy", "fz", "fghold"):
                out[n] = self.arrays[n].copy()
                continue
            api.sync_host(self.arrays[n])
            out[n] = self.arrays[n].copy()
        return out

    def press_stats(self):
        return api.press_stats(self.w.im, self.w.jm, self.w.km)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def ulp_diff(a, b):
    """max |a-b| in units of the larger operand's ulp (FP32)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(np.float32))
    sp = np.where(sp == 0, np.float32(1e-45), sp)
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / sp)) if a.size else 0.0
