"""les_press_noalloc (the press entry point that allocates nothing) against
les_press, the entry point it is a variant of, through the C ABI: two calls
in a row on one device, the second on a smaller domain, each from the state
a step leaves before its press.  p, rhs, f, g, h must equal les_press's bit
for bit (the grids have dx1 = dy1 = 4, so the FP64 volume sum is exact and
rhsav has the same bits), with the same SOR iteration count."""
import ctypes

import numpy as np
import pytest

from backends import GpuBackend
from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu

ARGS = api.SIGNATURES["press"]
OUT = ("rhs", "p", "f", "g", "h")


class Hip:
    """hipMalloc/hipMemcpy/hipFree through the HIP runtime the library uses:
    the caller's own device memory."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def put(self, a: np.ndarray) -> ctypes.c_void_p:
        a = np.asfortranarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 4))) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, ctypes.c_size_t(a.nbytes), 1) == 0  # host to device
        self.live.append(p)
        return p

    def get(self, p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(np.asfortranarray(like))
        assert self.rt.hipMemcpy(out.ctypes.data, p, ctypes.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _press_noalloc(hip, gpu, n, nmax):
    """les_press_noalloc on device copies of the backend's host arrays."""
    L = api.lib()
    f = L.les_press_noalloc
    f.argtypes = ([ctypes.c_int if a in api.INT_ARGS else ctypes.c_float if a in api.REAL_ARGS else ctypes.c_void_p
                   for a in ARGS] + [ctypes.c_void_p, ctypes.c_void_p])
    f.restype = ctypes.c_int
    dev = {}
    vals = []
    for a in ARGS:
        if a in ("n", "nmax"):
            vals.append(n if a == "n" else nmax)
        elif a in gpu.scalars:
            vals.append(gpu.scalars[a])
        else:
            dev[a] = hip.put(gpu.arrays[a])
            vals.append(dev[a])
    stats = hip.put(np.array([-7, -7], dtype=np.int32))
    rc = f(*vals, None, stats)
    assert rc == 0, L.les_last_error().decode()
    assert hip.rt.hipDeviceSynchronize() == 0
    out = {a: hip.get(dev[a], gpu.arrays[a]) for a in OUT}
    st = hip.get(stats, np.zeros(2, dtype=np.int32))
    return out, st


@pytest.mark.parametrize("dims", [[(48, 40, 24), (24, 20, 16)]], ids=["48x40x24-then-24x20x16"])
def test_noalloc_press_equals_les_press(dims):
    hip = Hip()
    try:
        for call, (im, jm, km) in enumerate(dims):
            w = setup_gen.small_workload(im, jm, km, seed=9 + call)
            s = setup_gen.build(w)
            print(f"\n[press noalloc] call {call + 1}: {im}x{jm}x{km}", flush=True)
            gpu = GpuBackend(s)
            gpu.step(1, 4)
            gpu.step(2, 4, stop_after="adam")
            state = gpu.fetch(["u", "v", "w", "p", "f", "g", "h", "rhs"])
            gpu.load(state)  # host arrays current; the next call uploads them again
            got, st = _press_noalloc(hip, gpu, 2, 4)
            gpu.call("press", n=2, nmax=4)
            want = gpu.fetch(list(OUT))
            for a in OUT:
                assert np.array_equal(got[a], want[a]), (call, a)
            ref = gpu.press_stats()
            assert st[0] == ref["iters"] and st[1] == 0, (st, ref["iters"])
            hip.free()
    finally:
        hip.free()
