"""The step's neighbours on the device: aveflow's time statistics
(aveflow.f95) and bondv1_data24's file-driven inflow (bondv1_data24.f95),
against the restatement (bit for bit: aveflow is adds, products, one
division and a square root per value, all correctly rounded in FP32), through
the reference-signature C ABI, the domain API (graph replay) and the
drop-in Fortran modules."""
import numpy as np
import pytest

import fortran_loop
from backends import GpuBackend
from lesamd import api, fortran_io, layout, setup_gen
from oracle import oracle

pytestmark = pytest.mark.gpu
CASE = dict(im=16, jm=12, km=10, seed=21)
STATE = ("u", "v", "w", "p", "sm")


def _setup():
    return setup_gen.build(setup_gen.small_workload(ustar=None, **CASE))


def _planes(w, n, rng=None):
    rng = rng or np.random.default_rng(100 + n)
    base = np.linspace(2.0, 6.0, w.km)[None, :] * np.ones((w.jm, 1))
    ua = (base * (1 + 0.05 * rng.uniform(-1, 1, base.shape))).astype(np.float32)
    va = (0.3 * rng.uniform(-1, 1, base.shape)).astype(np.float32)
    wa = (0.1 * rng.uniform(-1, 1, base.shape)).astype(np.float32)
    return ua, va, wa


def _assert_stats_equal(got, want, what):
    for n in layout.AVE_FIELDS:
        assert np.array_equal(got[n], want[n]), (what, n, float(np.abs(got[n] - want[n]).max()))


@pytest.mark.parametrize("n1,nmax", [(1, 3), (2, 4)])
def test_aveflow_reference_signature(n1, nmax):
    s = _setup()
    w = s.workload
    print(f"\n[aveflow] C ABI {w.im}x{w.jm}x{w.km}, n1={n1} nmax={nmax}", flush=True)
    g = GpuBackend(s)
    dev = {n: layout.new_field(n, w.im, w.jm, w.km) for n in layout.AVE_FIELDS}
    ref = {n: layout.new_field(n, w.im, w.jm, w.km) for n in layout.AVE_FIELDS}
    for n in range(1, nmax + 1):
        g.step(n, nmax)
        g.call("aveflow", n=n, n1=n1, nmax=nmax, **dev)
        fields = g.fetch(STATE)
        oracle.oracle_aveflow(n, n1, nmax, ref, fields, w.im, w.jm, w.km, prec=32)
    _assert_stats_equal(dev, ref, "reference signature")
    assert np.all(dev["avel"] == 0.0)
    assert np.all(dev["aveuu"] >= 0.0)


def test_aveflow_domain_api_graph():
    s = _setup()
    w = s.workload
    n1, nmax = 2, 5
    print(f"\n[aveflow] domain API + graph {w.im}x{w.jm}x{w.km}", flush=True)
    api.release_all()
    d = api.Domain(s)
    d.set_graph(True)
    ref = {n: layout.new_field(n, w.im, w.jm, w.km) for n in layout.AVE_FIELDS}
    for n in range(1, nmax + 1):
        d.step(n, nmax)
        d.aveflow(n, n1, nmax)
        fields = {f: d.download(f) for f in STATE}
        oracle.oracle_aveflow(n, n1, nmax, ref, fields, w.im, w.jm, w.km, prec=32)
    got = {n: d.download(n) for n in layout.AVE_FIELDS}
    d.close()
    _assert_stats_equal(got, ref, "domain")


@pytest.mark.parametrize("n", [1, 3])
def test_bondv1_data24_stage_matches_restatement(n):
    s = _setup()
    w = s.workload
    print(f"\n[bondv1_data24] stage n={n}", flush=True)
    rng = np.random.default_rng(n)
    ua, va, wa = _planes(w, n, rng)
    g = GpuBackend(s)
    o = oracle.OracleBackend(s, prec=32)
    init = {f: rng.uniform(-1, 1, g.arrays[f].shape).astype(np.float32) for f in ("u", "v", "w")}
    g.load(init)
    o.load(init)
    api.set_inflow_uvw(w.im, w.jm, w.km, ua, va, wa)
    g.call("bondv1", n=n)
    o.bondv1_data24(n, ua, va, wa)
    got = g.fetch(["u", "v", "w"])
    for f in ("u", "v", "w"):
        assert np.array_equal(got[f], o.arrays[f]), f


def test_domain_inflow_planes_between_graph_replays():
    s = _setup()
    w = s.workload
    nmax = 4
    print("\n[bondv1_data24] domain planes, graph vs eager", flush=True)
    api.release_all()
    out = []
    for graph in (False, True):
        d = api.Domain(s)
        d.set_graph(graph)
        for n in range(1, nmax + 1):
            d.set_inflow(*_planes(w, n))
            d.step(n, nmax)
        out.append({f: d.download(f) for f in STATE})
        d.close()
    for f in STATE:
        assert np.array_equal(out[0][f], out[1][f]), f
    ua, va, wa = _planes(w, nmax)
    # u,v (0:im+1,-1:jm+1,0:km+1): bondv1 of the last step set i=0,1 to its planes
    assert np.array_equal(out[1]["u"][0, 2:w.jm + 2, 1:w.km + 1], ua)
    assert np.array_equal(out[1]["v"][1, 2:w.jm + 2, 1:w.km + 1], va)


def test_fortran_dropin_data24_and_aveflow(tmp_path, monkeypatch):
    s = _setup()
    w = s.workload
    nmax, n1 = 3, 1
    print(f"\n[fortran drop-in] bondv1_data24 + aveflow, steps 1-{nmax}", flush=True)
    d24 = tmp_path / "d24"
    d24.mkdir()
    for n in range(1, nmax + 1):
        ua, va, wa = _planes(w, n)
        full = []
        for a in (ua, va, wa):
            b = np.zeros((fortran_io.DATA24_RECORD, w.km), np.float32)
            b[fortran_io.DATA24_OFFSET:fortran_io.DATA24_OFFSET + w.jm] = a
            full.append(b)
        fortran_io.write_data24(d24 / fortran_io.data24_name(n), *full)

    # the same steps from Python through the C ABI
    py = GpuBackend(s)
    want = {n: layout.new_field(n, w.im, w.jm, w.km) for n in layout.AVE_FIELDS}
    for n in range(1, nmax + 1):
        ua, va, wa = fortran_io.read_data24(d24 / fortran_io.data24_name(n), w.jm, w.km)
        api.set_inflow_uvw(w.im, w.jm, w.km, ua, va, wa)
        py.step(n, nmax)
        py.call("aveflow", n=n, n1=n1, nmax=nmax, **want)
    state = py.fetch(STATE)

    fo = GpuBackend(s)
    monkeypatch.chdir(tmp_path)
    got = fortran_loop.run(fo.arrays, fo.scalars, 1, nmax, nmax, "data20", "data21", data24_dir=str(d24),
                           nave1=n1, data10="data10", data11="data11")
    for f in STATE:
        assert np.array_equal(fo.arrays[f], state[f]), f
    _assert_stats_equal(got, want, "fortran")
    d10 = fortran_io.read_data10(tmp_path / "data10", w.im, w.jm, w.km)
    d11 = fortran_io.read_data11(tmp_path / "data11", w.im, w.jm, w.km)
    for n in fortran_io.DATA10:
        assert np.array_equal(d10[n], fortran_io.ave_interior(n, want[n], w.km)), n
    for n in fortran_io.DATA11:
        assert np.array_equal(d11[n], fortran_io.ave_interior(n, want[n], w.km)), n
