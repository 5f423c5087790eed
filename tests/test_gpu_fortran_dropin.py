"""The drop-in Fortran modules (mpi-les_amd/fortran/les_amd_modules.F90)
driven by the time loop of main.f95:247-296 (tests/fortran/time_loop.F90),
on the GPU:

* after two steps every state array equals, bit for bit, the same two steps
  run through the C ABI from Python (the modules pass the reference's
  arguments unchanged), and the reference's own outputs of those steps
  (tests/golden/stages_even.npz) to the per-step bound of 1e-5;
* the data20 and data21 dumps the modules write at n == nmax
  (press.f95:205-217, adam.f95:20-29) hold those arrays, big-endian.
"""
import numpy as np
import pytest

import fortran_loop
from backends import GpuBackend, rel_l2
from lesamd import fortran_io, layout, setup_gen

pytestmark = pytest.mark.gpu
CASE = dict(im=16, jm=12, km=10, seed=21)
FINAL = {"u": "2/bondv1/u", "v": "2/bondv1/v", "w": "2/bondv1/w", "p": "2/press/p", "rhs": "2/press/rhs",
         "f": "2/press/f", "g": "2/press/g", "h": "2/press/h", "usum": "2/feedbf/usum",
         "vsum": "2/feedbf/vsum", "wsum": "2/feedbf/wsum", "sm": "2/les/sm", "fold": "2/adam/fold",
         "gold": "2/adam/gold", "hold": "2/adam/hold", "uspd": "2/velfg/uspd", "vspd": "2/velfg/vspd"}


def test_fortran_time_loop_through_dropin_modules(tmp_path, monkeypatch):
    from pathlib import Path
    fx = np.load(Path(__file__).resolve().parent / "golden" / "stages_even.npz")
    s = setup_gen.build(setup_gen.small_workload(ustar=None, **CASE))
    w = s.workload
    print(f"\n[fortran drop-in] {w.im}x{w.jm}x{w.km}, steps 1-2", flush=True)

    # the same two steps through the C ABI from Python
    py = GpuBackend(s)
    fold1 = None
    for n in (1, 2):
        py.step(n, 2)
        if n == 1:
            fold1 = py.fetch(["fold", "gold", "hold"])
    want = py.fetch(list(FINAL))

    fo = GpuBackend(s)  # fresh host arrays in reference layout (releases the previous domain)
    monkeypatch.chdir(tmp_path)
    fortran_loop.run(fo.arrays, fo.scalars, 1, 2, 2, "data20", "data21")
    for k in FINAL:
        assert np.array_equal(fo.arrays[k], want[k]), k
        assert rel_l2(fo.arrays[k], fx[FINAL[k]]) <= 1e-5, (k, rel_l2(fo.arrays[k], fx[FINAL[k]]))

    d20 = fortran_io.read_data20(tmp_path / "data20", w.im, w.jm, w.km)
    for k in fortran_io.DATA20:
        assert np.array_equal(d20[k], layout.interior(k, fo.arrays[k], w.im, w.jm, w.km)), k
    d21 = fortran_io.read_data21(tmp_path / "data21", w.im, w.jm, w.km)
    for k in ("fold", "gold", "hold"):
        assert np.array_equal(d21[k], layout.interior(k, fold1[k], w.im, w.jm, w.km)), k
    assert not d21["fghold"].any()
