"""Fortran-side pieces that need no GPU: the unformatted record format of
the data20/data21 dumps, and the drop-in modules' test library (built by
mpi-les_amd/build.py with flang) loading with its entry point."""
import ctypes

import numpy as np
import pytest

import fortran_loop
from lesamd import fortran_io


def test_record_layout_is_big_endian_with_length_markers(tmp_path):
    p = tmp_path / "r.bin"
    fortran_io.write_records(p, [np.array([1.0, 2.0, 3.0], ">f4").tobytes()])
    raw = p.read_bytes()
    # a 12-byte record: marker 12 (big-endian), 1.0f = 3f800000, ..., marker 12
    assert raw[:8] == bytes([0, 0, 0, 12, 0x3F, 0x80, 0, 0])
    assert raw[-4:] == bytes([0, 0, 0, 12]) and len(raw) == 20
    assert fortran_io.read_records(p) == [raw[4:16]]


def test_data20_round_trip_and_truncation(tmp_path):
    rng = np.random.default_rng(0)
    im, jm, km = 5, 4, 3
    fields = {k: rng.standard_normal((im, jm, km)).astype(np.float32) for k in fortran_io.DATA20}
    p = tmp_path / "data20"
    fortran_io.write_fields(p, fields, fortran_io.DATA20)
    got = fortran_io.read_data20(p, im, jm, km)
    for k in fields:
        assert np.array_equal(got[k], fields[k])
    p.write_bytes(p.read_bytes()[:-3])
    with pytest.raises(ValueError):
        fortran_io.read_data20(p, im, jm, km)
    with pytest.raises(ValueError):
        fortran_io.read_data20(tmp_path / "data20", im, jm, km + 1)


def test_dropin_module_library_exports_time_loop():
    if not fortran_loop.LIB.exists():
        pytest.skip("tests/fortran/liblestimeloop.so not built (flang absent)")
    lib = ctypes.CDLL(str(fortran_loop.LIB))
    assert hasattr(lib, "les_test_time_loop")


def test_restart_pair_layout_and_round_trip(tmp_path):
    """data30/data31 as anime.f95:326-595 writes them: an 8-byte (n, time)
    record, then one interior record per field, big-endian."""
    from lesamd import layout, restart
    im, jm, km = 6, 5, 4
    rng = np.random.default_rng(1)
    fields = {}
    for nm in restart.DATA30 + restart.DATA31:
        a = layout.new_field(nm, im, jm, km)
        a[...] = rng.standard_normal(a.shape).astype(np.float32)
        fields[nm] = a
    p30, p31 = restart.write_restart(tmp_path, 48000, 2399.95, fields, im, jm, km)
    assert p30.name == "data30048000.dat" and p31.name == "data31048000.dat"
    recs = fortran_io.read_records(p30)
    assert len(recs) == 8 and len(recs[0]) == 8 and len(recs[1]) == 4 * im * jm * km
    assert recs[0][:4] == (48000).to_bytes(4, "big")
    assert len(fortran_io.read_records(p31)) == 6
    n, t, got = restart.read_restart(p30, p31, im, jm, km)
    assert n == 48000 and abs(t - 2399.95) < 1e-3
    for nm in got:
        assert np.array_equal(got[nm], layout.interior(nm, fields[nm], im, jm, km)), nm
    with pytest.raises(ValueError):
        restart.read_restart(p31, p30, im, jm, km)
