"""ctypes access to the two CPU checkers -- TEST INFRASTRUCTURE ONLY.

* ``liblesoracle.so``: the C restatement (les_oracle.c), FP32 (orc32_*) and
  FP64 (orc64_*) instantiations.
* ``_ref/<variant>/libref.so``: the reference's own Fortran subroutines,
  compiled unchanged with flang by oracle/Makefile (``make ref``), behind
  the bind(C) wrappers of ref_harness.f90.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
use this module.  Each backend object holds the arrays of one domain in
the reference layout and calls the stages with the reference's argument
lists (main.f95:260-296).
"""
from __future__ import annotations

import ctypes
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(ROOT / "mpi-les_amd"))
from lesamd import layout  # noqa: E402

INT_ARGS = {"km", "jm", "im", "n", "nmax", "n1"}
REAL_ARGS = {"ro", "dt", "vn", "alpha", "beta"}

# Argument lists of the reference subroutines (velnw.f95:5, bondv1.f95:5,
# velFG.f95:6-8, feedbf.f95:5-6, les.f95:6-7, adam.f95:5, press.f95:9-10,
# bondFG.f95:5, boundp.f95:7,28, boundsm.f95:5), minus the character
# file-name arguments of adam and press.
SIGNATURES = {
    "velnw": "km jm im p ro dxs u dt f dys v g dzs w h",
    "bondv1": "jm u z2 dzn v w km n im dt dxs",
    "velfg": "km jm im dx1 cov1 cov2 cov3 dfu1 diu1 diu2 dy1 diu3 dzn vn f cov4 cov5 cov6 dfv1 diu4 "
             "diu5 diu6 g cov7 cov8 cov9 dfw1 diu7 diu8 diu9 dzs h nou1 u nou5 v nou9 w nou2 nou3 nou4 "
             "nou6 nou7 nou8 uspd vspd",
    "feedbf": "km jm im usum u bmask1 vsum v cmask1 wsum w dmask1 alpha dt beta fx fy fz f g h",
    "les": "km delx1 dx1 dy1 dzn jm im diu1 diu2 diu3 diu4 diu5 diu6 diu7 diu8 diu9 sm f g h uspd vspd dxs dys",
    "adam": "n nmax fold im jm km gold hold fghold f g h",
    "press": "km jm im rhs u dx1 v dy1 w dzn f g h dt cn1 cn2l p cn2s cn3l cn3s cn4l cn4s n nmax usum vsum wsum",
    "bondfg": "km jm f im g h",
    "boundp1": "km jm p im",
    "boundp2": "jm im p km",
    "boundsm": "km jm sm im",
}
SIGNATURES = {k: v.split() for k, v in SIGNATURES.items()}
# Restated only (the reference's versions need MPI): bondv1_data24.f95:6 with
# the planes its file records hold as ua, va, wa, and aveflow.f95:5-6 minus
# the arguments its statistics never touch.
ORACLE_ONLY = {
    "bondv1_data24": "jm u z2 dzn v w km n im dt dxs ua va wa".split(),
    "aveflow": ("n n1 km jm im aveu avev avew avep avel aveuu avevv aveww avesm avesmsm uwfx "
                "u v w p sm nmax").split(),
}
STAGES = ("velnw", "bondv1", "velfg", "feedbf", "les", "adam", "press")


def _argtypes(names, real_ct):
    out = []
    for a in names:
        if a in INT_ARGS:
            out.append(ctypes.c_int)
        elif a in REAL_ARGS:
            out.append(real_ct)
        else:
            out.append(ctypes.c_void_p)
    return out


def oracle_lib_path(omp: bool = False) -> Path:
    """The restatement; omp: its OpenMP build (CPU timing baseline only)."""
    return HERE / ("liblesoracle_omp.so" if omp else "liblesoracle.so")


def ref_lib_path(variant: str) -> Path:
    return HERE / "_ref" / variant / "libref.so"


def ref_available(variant: str = "small") -> bool:
    return ref_lib_path(variant).exists()


class _Domain:
    """Arrays of one domain + the scalars the stages take."""

    def __init__(self, setup, dtype, ext=None):
        w = setup.workload
        self.w = w
        self.im, self.jm, self.km = w.im, w.jm, w.km
        self.ext = ext or (w.im, w.jm, w.km)      # allocated (compile-time) extents
        ip, jp, kp = self.ext
        self.dtype = dtype
        self.arrays = {}
        for n in layout.STEP_FIELDS:
            self.arrays[n] = layout.new_field(n, ip, jp, kp, dtype)
        for n, a in setup.masks.items():
            self.arrays[n] = self._embed(n, a.astype(dtype))
        for n, a in setup.metrics.items():
            self.arrays[n] = self._embed(n, a.astype(dtype))
        for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s"):
            self.arrays[n] = self._embed(n, setup.coeffs[n].astype(dtype))
        cn1 = layout.new_field("cn1", ip, jp, kp, dtype)
        cn1[: w.im, : w.jm, : w.km] = setup.coeffs["cn1"]
        self.arrays["cn1"] = cn1
        self.arrays["delx1"] = layout.new_metric("delx1", ip, jp, kp, dtype)
        self.scalars = dict(km=w.km, jm=w.jm, im=w.im, ro=1.1763, dt=w.dt, alpha=-10.0, beta=-1.0)
        from lesamd.setup_gen import VN
        self.scalars["vn"] = float(VN)

    def _embed(self, name, small):
        ip, jp, kp = self.ext
        if small.ndim == 1:
            big = layout.new_metric(name, ip, jp, kp, small.dtype)
        else:
            big = layout.new_field(name, ip, jp, kp, small.dtype)
        layout.embed(name, small, big, self.im, self.jm, self.km)
        return big

    def load(self, state: dict):
        for n, a in state.items():
            dst = self.arrays[n]
            sl = tuple(slice(0, s) for s in a.shape)
            dst[sl] = a.astype(self.dtype)

    def fetch(self, names):
        return {n: layout.extract(n, self.arrays[n], self.im, self.jm, self.km) for n in names}

    def args(self, stage, **extra):
        vals = []
        for a in SIGNATURES[stage] if stage in SIGNATURES else ORACLE_ONLY[stage]:
            if a in extra:
                vals.append(extra[a])
            elif a in self.scalars:
                vals.append(self.scalars[a])
            else:
                vals.append(self.arrays[a].ctypes.data)
        return vals


class OracleBackend:
    """The C restatement.  prec=32 or 64; redmode 0 = sequential FP32
    reductions (the shipped build), 1 = FP64 accumulation (the GPU path's
    arithmetic).  omp: the OpenMP build (k loops over threads; only for the
    CPU timing baseline, with redmode 1)."""

    def __init__(self, setup, prec=32, redmode=1, omp=False):
        self.lib = ctypes.CDLL(str(oracle_lib_path(omp)))
        self.prefix = "orc32_" if prec == 32 else "orc64_"
        real_ct = ctypes.c_float if prec == 32 else ctypes.c_double
        self.dom = _Domain(setup, np.float32 if prec == 32 else np.float64)
        self.setup = setup
        self.fn = {}
        for s, names in list(SIGNATURES.items()) + list(ORACLE_ONLY.items()):
            f = getattr(self.lib, self.prefix + s)
            f.argtypes = _argtypes(names, real_ct)
            f.restype = None
            self.fn[s] = f
        w = setup.workload
        cfg = getattr(self.lib, self.prefix + "press_config")
        cfg.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        cfg(w.nmaxp, w.pjuge, w.omega, redmode)
        self._stats = getattr(self.lib, self.prefix + "press_stats")
        self._stats.argtypes = [ctypes.c_void_p] * 7
        self._set_inflow = getattr(self.lib, self.prefix + "bondv1_set_inflow")
        self._set_inflow.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        self._prof = None
        if setup.inflow is not None:
            self._prof = np.asfortranarray(setup.inflow.astype(self.dom.dtype))
            self._set_inflow(self._prof.ctypes.data, w.jm, w.km)
        else:
            self._set_inflow(None, 0, 0)

    @property
    def arrays(self):
        return self.dom.arrays

    def call(self, stage, **extra):
        self.fn[stage](*self.dom.args(stage, **extra))

    def step(self, n, nmax, stop_after=None):
        for s in STAGES:
            kw = {"n": n, "nmax": nmax} if s in ("bondv1", "adam", "press") else {}
            self.call(s, **kw)
            if s == stop_after:
                return

    def bondv1_data24(self, n, ua, va, wa):
        """bondv1 with inflow planes (jm, km) for u, v, w (bondv1_data24.f95)."""
        planes = [np.asfortranarray(a, dtype=self.dom.dtype) for a in (ua, va, wa)]
        self.call("bondv1_data24", n=n, ua=planes[0].ctypes.data, va=planes[1].ctypes.data,
                  wa=planes[2].ctypes.data)

    def press_stats(self, hist_len=4096):
        it = ctypes.c_int()
        hist = np.zeros(hist_len)
        vals = [ctypes.c_double() for _ in range(4)]
        self._stats(ctypes.addressof(it), hist.ctypes.data, hist_len, *[ctypes.addressof(v) for v in vals])
        return dict(iters=it.value, sor=hist[: it.value].copy(), rhsav=vals[0].value, area=vals[1].value,
                    pav=vals[2].value, pco=vals[3].value)

    def fetch(self, names):
        return self.dom.fetch(names)

    def load(self, state):
        self.dom.load(state)


class RefBackend:
    """The reference's own subroutines (flang build).  Arrays are allocated
    at the build's compile-time extents and the domain's index range is
    embedded in them; the reference's inflow is its fixed 5 m/s.  The o64
    variants (kind-promoted FP64 builds) take and return FP64 arrays."""

    def __init__(self, setup, variant="small"):
        path = ref_lib_path(variant)
        self.lib = ctypes.CDLL(str(path))
        dims = [ctypes.c_int() for _ in range(5)]
        self.lib.ref_dims(*[ctypes.byref(d) for d in dims])
        ip, jp, kp, self.ipmax, self.jpmax = [d.value for d in dims]
        w = setup.workload
        if w.im > ip or w.jm > jp or w.km > kp:
            raise ValueError(f"domain {w.im}x{w.jm}x{w.km} exceeds the {variant} build ({ip}x{jp}x{kp})")
        if setup.inflow is not None:
            raise ValueError("the reference build only has its fixed 5 m/s inflow (bondv1.f95:33-54)")
        self.fp64 = variant.startswith("o64")
        self.dom = _Domain(setup, np.float64 if self.fp64 else np.float32, ext=(ip, jp, kp))
        self.setup = setup
        self.fn = {}
        for s, names in SIGNATURES.items():
            f = getattr(self.lib, "ref_" + s)
            f.argtypes = _argtypes(names, ctypes.c_double if self.fp64 else ctypes.c_float)
            f.restype = None
            self.fn[s] = f

    @property
    def arrays(self):
        return self.dom.arrays

    def flush(self):
        """Flush the reference's unit-6 output (its progress prints)."""
        if hasattr(self.lib, "ref_flush"):
            self.lib.ref_flush()

    def call(self, stage, **extra):
        self.fn[stage](*self.dom.args(stage, **extra))

    def step(self, n, nmax, stop_after=None):
        for s in STAGES:
            kw = {"n": n, "nmax": nmax} if s in ("bondv1", "adam", "press") else {}
            self.call(s, **kw)
            if s == stop_after:
                return

    def fetch(self, names):
        return self.dom.fetch(names)

    def load(self, state):
        self.dom.load(state)

    # setup routines of the reference, for pinning the generator
    def ref_grid(self):
        ip, jp, kp = self.dom.ext
        names = ("dx1", "dxl", "dy1", "dyl", "z2", "dzn", "dzs", "dxs", "dys")
        arrs = {n: layout.new_metric(n, ip, jp, kp) for n in names}
        self.lib.ref_grid(*[ctypes.c_void_p(arrs[n].ctypes.data) for n in names])
        return arrs

    def ref_set(self):
        ints = [ctypes.c_int() for _ in range(6)]
        reals = [ctypes.c_float() for _ in range(5)]
        self.lib.ref_set(*[ctypes.byref(x) for x in ints + reals])
        keys = ("im", "jm", "km", "n0", "n1", "nmax", "dt", "ro", "vn", "alpha", "beta")
        return dict(zip(keys, [x.value for x in ints + reals]))

    def ref_init(self, metrics):
        """Run init.f95 on the given metrics; returns cn* and zeroed fields."""
        ip, jp, kp = self.dom.ext
        w = self.setup.workload
        a = {n: layout.new_field(n, ip, jp, kp) for n in ("u", "v", "w", "p", "amask1", "bmask1", "cmask1", "dmask1", "cn1")}
        for n in ("u", "v", "w", "p"):
            a[n][:] = 7.0  # init must zero these
        c = {n: layout.new_metric(n, ip, jp, kp) for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s")}
        m = {n: self.dom._embed(n, metrics[n]) for n in ("dxs", "dys", "dzs", "z2", "dzn")}
        zbm = np.zeros((self.ipmax + 3, self.jpmax + 3), dtype=np.float32, order="F")
        order = ["u", "v", "w", "p", "cn2s", "dxs", "cn2l", "cn3s", "dys", "cn3l", "dzs", "cn4s", "cn4l", "cn1",
                 "amask1", "bmask1", "cmask1", "dmask1", "zbm", "z2", "dzn"]
        allv = {**a, **c, **m, "zbm": zbm}
        f = self.lib.ref_init
        f.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * len(order)
        f(w.km, w.jm, w.im, *[allv[n].ctypes.data for n in order])
        out = {n: layout.extract(n, c[n], w.im, w.jm, w.km) for n in c}
        out["cn1"] = np.asfortranarray(a["cn1"][: w.im, : w.jm, : w.km])
        for n in ("u", "v", "w", "p"):
            out[n] = layout.extract(n, a[n], w.im, w.jm, w.km)
        return out


def oracle_aveflow(n: int, n1: int, nmax: int, stats: dict, fields: dict, im: int, jm: int, km: int,
                   prec: int = 32):
    """aveflow (aveflow.f95:80-203) of the restatement on host arrays in
    reference layout: `stats` (layout.AVE_FIELDS, updated in place) and
    `fields` u, v, w, p, sm (read)."""
    lib = ctypes.CDLL(str(oracle_lib_path()))
    f = getattr(lib, ("orc32_" if prec == 32 else "orc64_") + "aveflow")
    f.argtypes = _argtypes(ORACLE_ONLY["aveflow"], ctypes.c_float)
    f.restype = None
    dt = np.float32 if prec == 32 else np.float64
    for name, a in stats.items():  # updated in place
        if a.dtype != dt or not a.flags.f_contiguous:
            raise ValueError(f"{name}: need a Fortran-ordered {dt.__name__} array")
    fields = {name: np.asfortranarray(a, dtype=dt) for name, a in fields.items()}
    vals = []
    for a in ORACLE_ONLY["aveflow"]:
        if a in INT_ARGS:
            vals.append(dict(n=n, n1=n1, km=km, jm=jm, im=im, nmax=nmax)[a])
        else:
            vals.append((stats[a] if a in stats else fields[a]).ctypes.data)
    f(*vals)
