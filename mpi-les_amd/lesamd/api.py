"""Python host interface over the C ABI of liblesamd.so (include/les_amd.h).

Two layers, as in the C ABI:

* module functions ``velnw, bondv1, velfg, feedbf, les, adam, press`` take
  the reference subroutines' argument lists (velnw.f95:5, bondv1.f95:5,
  velFG.f95:6-8, feedbf.f95:5-6, les.f95:6-7, adam.f95:5, press.f95:9-10,
  without the character file-name arguments) with numpy arrays in the
  reference layout (see layout.py); the library mirrors each array on the
  device and the host copy is stale until ``sync_host``;
* ``Domain`` owns a device-resident domain built from a set-up
  (setup_gen.DomainSetup) and advances it by whole steps.

The library must be built (mpi-les_amd/build.py); there is no fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from . import layout

# LESAMD_LIBRARY selects another build of the library (same ABI)
LIB_PATH = Path(os.environ.get("LESAMD_LIBRARY") or Path(__file__).with_name("liblesamd.so"))
_lib = None

INT_ARGS = {"km", "jm", "im", "n", "nmax", "n1"}
REAL_ARGS = {"ro", "dt", "vn", "alpha", "beta"}

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
    # aveflow.f95:5-6 without the arguments its statistics never touch
    "aveflow": "n n1 km jm im aveu avev avew avep avel aveuu avevv aveww avesm avesmsm uwfx u v w p sm nmax",
}
SIGNATURES = {k: v.split() for k, v in SIGNATURES.items()}
STAGES = ("velnw", "bondv1", "velfg", "feedbf", "les", "adam", "press")


class LesError(RuntimeError):
    pass


class DomainDesc(ctypes.Structure):
    _fields_ = [("im", ctypes.c_int), ("jm", ctypes.c_int), ("km", ctypes.c_int)] + [
        (n, ctypes.c_void_p) for n in ("dx1", "dy1", "dzn", "dzs", "dxs", "dys", "cn2l", "cn2s", "cn3l",
                                       "cn3s", "cn4l", "cn4s", "cn1", "bmask1", "cmask1", "dmask1", "inflow")
    ] + [(n, ctypes.c_float) for n in ("ro", "vn", "alpha", "beta", "dt")] + [
        ("nmaxp", ctypes.c_int), ("pjuge", ctypes.c_float), ("omega", ctypes.c_float),
        ("zbm", ctypes.c_void_p), ("z2", ctypes.c_void_p)]


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LesError(f"{LIB_PATH} is missing: build it with `python mpi-les_amd/build.py`")
    L = ctypes.CDLL(str(LIB_PATH))
    L.les_last_error.restype = ctypes.c_char_p
    for s, names in SIGNATURES.items():
        f = getattr(L, "les_" + s)
        f.argtypes = [ctypes.c_int if a in INT_ARGS else ctypes.c_float if a in REAL_ARGS else ctypes.c_void_p
                      for a in names]
        f.restype = ctypes.c_int
    vp, ci, cf, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
    L.les_bondv1_set_inflow.argtypes = [ci, ci, ci, vp]
    L.les_bondv1_set_inflow_uvw.argtypes = [ci, ci, ci, vp, vp, vp]
    L.les_domain_aveflow.argtypes = [vp, ci, ci, ci]
    L.les_domain_set_inflow.argtypes = [vp, vp, vp, vp]
    L.les_press_config.argtypes = [ci, cf, cf]
    L.les_press_stats.argtypes = [ci, ci, ci, vp, vp, ci, vp, vp, vp]
    L.les_sync_host.argtypes = [vp]
    L.les_mark_host_dirty.argtypes = [vp]
    L.les_domain_create.argtypes = [ctypes.POINTER(DomainDesc), ctypes.POINTER(vp)]
    L.les_domain_destroy.argtypes = [vp]
    L.les_domain_step.argtypes = [vp, ci, ci]
    L.les_domain_run.argtypes = [vp, ci, ci, ci]
    L.les_domain_stage.argtypes = [vp, ci, ci, ci]
    L.les_domain_upload.argtypes = [vp, ctypes.c_char_p, vp]
    L.les_domain_download.argtypes = [vp, ctypes.c_char_p, vp]
    L.les_domain_synchronize.argtypes = [vp]
    L.les_domain_query.argtypes = [vp]
    L.les_domain_set_press_mode.argtypes = [vp, ci, ci, ci]
    L.les_domain_set_press_omega.argtypes = [vp, cf]
    L.les_domain_bytes_needed.argtypes = [ci, ci, ci]
    L.les_domain_bytes_needed.restype = ctypes.c_longlong
    L.les_domain_press_stats.argtypes = [vp, vp, vp, ci, vp, vp, vp]
    L.les_domain_bytes.argtypes = [vp]
    L.les_domain_bytes.restype = ctypes.c_longlong
    L.les_domain_set_timing.argtypes = [vp, ci]
    L.les_domain_set_timing_runs.argtypes = [vp, ci]
    L.les_domain_set_graph.argtypes = [vp, ci]
    L.les_domain_autotune.argtypes = [vp, ci, vp, vp, vp, vp, vp, vp]
    L.les_domain_time_sor.argtypes = [vp, ci, vp]
    L.les_domain_stencil_chunks.argtypes = [vp, vp, vp, vp, vp, vp]
    L.les_domain_timing.argtypes = [vp, vp, vp]
    L.les_batch_create.argtypes = [vp, ci, ctypes.POINTER(vp)]
    L.les_batch_destroy.argtypes = [vp]
    L.les_batch_set_graph.argtypes = [vp, ci]
    L.les_batch_step.argtypes = [vp, vp, vp]
    L.les_batch_query.argtypes = [vp]
    L.les_batch_synchronize.argtypes = [vp]
    L.les_kernel_class_name.restype = ctypes.c_char_p
    L.les_kernel_class_name.argtypes = [ci]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise LesError(lib().les_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if a.dtype != np.float32 or not (a.flags.f_contiguous or a.ndim == 1):
            raise LesError("arrays must be float32 in Fortran order")
        return a.ctypes.data
    return a


def call(stage: str, **args):
    """Call a reference-signature entry point with named arguments."""
    vals = [args[a] if (a in INT_ARGS or a in REAL_ARGS) else _ptr(args[a]) for a in SIGNATURES[stage]]
    _check(getattr(lib(), "les_" + stage)(*vals))


def sync_host(arr: np.ndarray) -> bool:
    """Copy the device state back into arr; False if the library never saw
    arr (its host contents are then current)."""
    rc = lib().les_sync_host(arr.ctypes.data)
    if rc == 2:
        return False
    _check(rc)
    return True


def mark_host_dirty(arr: np.ndarray):
    _check(lib().les_mark_host_dirty(arr.ctypes.data))


def synchronize():
    _check(lib().les_synchronize())


def release_all():
    _check(lib().les_release_all())


def set_inflow(im, jm, km, prof):
    p = None if prof is None else np.asfortranarray(prof, dtype=np.float32)
    _check(lib().les_bondv1_set_inflow(im, jm, km, None if p is None else p.ctypes.data))
    return p


def set_inflow_uvw(im, jm, km, u_in, v_in=None, w_in=None):
    """Inflow planes (j,k) of bondv1_data24 for the next bondv1 calls on a
    domain of these extents; returns the arrays handed over."""
    planes = [None if a is None else np.asfortranarray(a, dtype=np.float32) for a in (u_in, v_in, w_in)]
    for a in planes:
        if a is not None and a.shape != (jm, km):
            raise LesError(f"inflow plane of shape {a.shape}, expected {(jm, km)}")
    _check(lib().les_bondv1_set_inflow_uvw(im, jm, km, *[None if a is None else a.ctypes.data for a in planes]))
    return planes


def press_config(nmaxp=50, pjuge=1e-4, omega=1.0):
    _check(lib().les_press_config(nmaxp, pjuge, omega))


def _stats(fn, *lead, hist_len=1024):
    it = ctypes.c_int()
    hist = np.zeros(hist_len)
    rh, pa = ctypes.c_double(), ctypes.c_double()
    nf = ctypes.c_int()
    _check(fn(*lead, ctypes.addressof(it), hist.ctypes.data, hist_len, ctypes.addressof(rh),
              ctypes.addressof(pa), ctypes.addressof(nf)))
    return dict(iters=it.value, sor=hist[: it.value].copy(), rhsav=rh.value, pav=pa.value, nonfinite=nf.value)


def press_stats(im, jm, km):
    return _stats(lib().les_press_stats, im, jm, km)


DOMAIN_FIELDS = ("u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "fold", "gold", "hold", "rhs",
                 "sm", "uspd", "vspd") + tuple(f"diu{q}" for q in range(1, 10)) + ("cn1", "bmask1", "cmask1", "dmask1")


class Domain:
    """A device-resident domain (les_domain_*)."""

    def __init__(self, setup, ro=1.1763, vn=None, alpha=-10.0, beta=-1.0):
        from .setup_gen import VN
        w = setup.workload
        self.workload = w
        self._keep = []

        def arr(a):
            a = np.ascontiguousarray(a, dtype=np.float32) if a.ndim == 1 else np.asfortranarray(a, dtype=np.float32)
            self._keep.append(a)
            return a.ctypes.data

        m, c = setup.metrics, setup.coeffs
        d = DomainDesc()
        d.im, d.jm, d.km = w.im, w.jm, w.km
        for n in ("dx1", "dy1", "dzn", "dzs", "dxs", "dys"):
            setattr(d, n, arr(m[n]))
        for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s"):
            setattr(d, n, arr(c[n]))
        # without a dense cn1 the library uses its separable definition
        d.cn1 = arr(c["cn1"]) if "cn1" in c else None
        if setup.masks is not None:
            for n in ("bmask1", "cmask1", "dmask1"):
                setattr(d, n, arr(setup.masks[n]))
        else:  # masks built on the device from the height map
            d.zbm = arr(setup.zbm)
            d.z2 = arr(m["z2"])
        d.inflow = arr(setup.inflow) if setup.inflow is not None else None
        d.ro, d.vn, d.alpha, d.beta, d.dt = ro, float(VN if vn is None else vn), alpha, beta, w.dt
        d.nmaxp, d.pjuge, d.omega = w.nmaxp, w.pjuge, w.omega
        h = ctypes.c_void_p()
        _check(lib().les_domain_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._keep = []

    @property
    def handle(self):
        return self._h

    def step(self, n, nmax):
        _check(lib().les_domain_step(self._h, n, nmax))

    def run(self, n0, n1, nmax):
        _check(lib().les_domain_run(self._h, n0, n1, nmax))

    def stage(self, s, n, nmax):
        _check(lib().les_domain_stage(self._h, STAGES.index(s) if isinstance(s, str) else s, n, nmax))

    def upload(self, name, arr):
        a = np.asfortranarray(arr, dtype=np.float32)
        _check(lib().les_domain_upload(self._h, name.encode(), a.ctypes.data))

    def download(self, name):
        w = self.workload
        out = layout.new_field(name, w.im, w.jm, w.km)
        _check(lib().les_domain_download(self._h, name.encode(), out.ctypes.data))
        return out

    def synchronize(self):
        _check(lib().les_domain_synchronize(self._h))

    def aveflow(self, n: int, n1: int, nmax: int):
        """aveflow after step n on the device (zero at n1, accumulate from
        n1, finalise at nmax); fetch with download("aveu") etc."""
        _check(lib().les_domain_aveflow(self._h, n, n1, nmax))

    def set_inflow(self, u_in, v_in=None, w_in=None):
        """Inflow planes (jm, km) for the next steps (bondv1_data24's u, v, w)."""
        w = self.workload
        planes = [None if a is None else np.asfortranarray(a, dtype=np.float32) for a in (u_in, v_in, w_in)]
        for a in planes:
            if a is not None and a.shape != (w.jm, w.km):
                raise LesError(f"inflow plane of shape {a.shape}, expected {(w.jm, w.km)}")
        _check(lib().les_domain_set_inflow(self._h, *[None if a is None else a.ctypes.data for a in planes]))

    def set_press_mode(self, mode: int, outer: int = 4, inner: int = 25):
        """0: the reference's SOR; 1: FP64 defect correction (`outer` steps
        of `inner` FP32 SOR iterations)."""
        _check(lib().les_domain_set_press_mode(self._h, mode, outer, inner))

    def set_press_omega(self, omega: float):
        """Relaxation of press mode 1's inner SOR iterations."""
        _check(lib().les_domain_set_press_omega(self._h, float(omega)))

    def done(self) -> bool:
        """True when every call issued on the domain has completed (no wait)."""
        r = lib().les_domain_query(self._h)
        if r < 0:
            raise LesError(lib().les_last_error().decode())
        return r == 1

    def press_stats(self):
        return _stats(lib().les_domain_press_stats, self._h)

    def nonfinite(self) -> bool:
        """True when the last pressure solve met a non-finite residual (the
        state has diverged).  Waits for the domain's stream."""
        return self.press_stats()["nonfinite"] != 0

    def autotune(self, reps: int = 0) -> dict:
        """Pick the SOR launch shape and the stencil kernels' k levels per
        thread for this grid shape (cached per shape; reps > 0 times them
        again with that many launches per candidate)."""
        r, ty, tw, kc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ub, ud = ctypes.c_double(), ctypes.c_double()
        _check(lib().les_domain_autotune(self._h, reps, ctypes.byref(r), ctypes.byref(ty), ctypes.byref(tw),
                                         ctypes.byref(kc), ctypes.byref(ub), ctypes.byref(ud)))
        out = dict(rows=r.value, tile_rows=ty.value, tile_width=tw.value, kchunk=kc.value, us_per_iter=ub.value,
                   us_per_iter_default=ud.value)
        out.update(self.stencil_chunks())
        return out

    def stencil_chunks(self) -> dict:
        """k levels per thread of velfg+feedbf, les_sm and les viscous+adam
        (0: automatic) and the tuner's timings of one velfg+feedbf and one
        les+adam launch sequence, tuned and automatic (us)."""
        v, sm, vc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ub, ud = ctypes.c_double(), ctypes.c_double()
        _check(lib().les_domain_stencil_chunks(self._h, ctypes.byref(v), ctypes.byref(sm), ctypes.byref(vc),
                                               ctypes.byref(ub), ctypes.byref(ud)))
        return dict(velfg_kchunk=v.value, sm_kchunk=sm.value, visc_kchunk=vc.value, stencil_us=ub.value,
                    stencil_us_default=ud.value)

    def time_sor(self, iters: int = 200) -> float:
        """Microseconds per fused SOR iteration of the current launch shape:
        `iters` back-to-back launches between one pair of HIP events."""
        us = ctypes.c_double()
        _check(lib().les_domain_time_sor(self._h, iters, ctypes.byref(us)))
        return us.value

    def set_graph(self, on: bool):
        """Replay steps n > 1 from one captured HIP graph."""
        _check(lib().les_domain_set_graph(self._h, 1 if on else 0))

    def set_timing(self, on: bool, classes=None, runs: bool = False):
        """Bracket launches with HIP events; classes: names to record (all if
        None); runs: one event pair per run of back-to-back launches of a class."""
        mask = 0
        if on:
            if classes is None:
                mask = -1
            else:
                L = lib()
                names = [L.les_kernel_class_name(c).decode() for c in range(L.les_kernel_classes())]
                for c in classes:
                    mask |= 1 << names.index(c)
        _check(lib().les_domain_set_timing(self._h, mask))
        if on:
            _check(lib().les_domain_set_timing_runs(self._h, 1 if runs else 0))

    def timing(self):
        """{kernel class: (total ms, launches)} since the previous read."""
        L = lib()
        n = L.les_kernel_classes()
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        _check(L.les_domain_timing(self._h, ms.ctypes.data, cnt.ctypes.data))
        return {L.les_kernel_class_name(c).decode(): (float(ms[c]), int(cnt[c])) for c in range(n)}

    def nbytes(self):
        return lib().les_domain_bytes(self._h)

    def close(self):
        if self._h:
            lib().les_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Domains advanced together (les_batch_*): one step of each per call,
    every stage one launch for all of them and the SOR iterations one launch
    per iteration."""

    def __init__(self, domains, graph: bool = True):
        self.domains = list(domains)
        hs = (ctypes.c_void_p * len(self.domains))(*[d.handle.value for d in self.domains])
        h = ctypes.c_void_p()
        _check(lib().les_batch_create(hs, len(self.domains), ctypes.byref(h)))
        self._h = h
        _check(lib().les_batch_set_graph(h, 1 if graph else 0))

    def step(self, ns, nmaxs):
        n = np.ascontiguousarray(ns, dtype=np.int32)
        m = np.ascontiguousarray(nmaxs, dtype=np.int32)
        if n.size != len(self.domains) or m.size != len(self.domains):
            raise LesError("one n and one nmax per domain")
        _check(lib().les_batch_step(self._h, n.ctypes.data, m.ctypes.data))

    def done(self) -> bool:
        r = lib().les_batch_query(self._h)
        if r < 0:
            raise LesError(lib().les_last_error().decode())
        return r == 1

    def synchronize(self):
        _check(lib().les_batch_synchronize(self._h))

    def close(self):
        if self._h:
            lib().les_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bytes_needed(im: int, jm: int, km: int) -> int:
    """Device memory a domain of these extents holds (for admission control)."""
    return int(lib().les_domain_bytes_needed(im, jm, km))
