"""The restatement's aveflow (aveflow.f95) and bondv1_data24
(bondv1_data24.f95), and the file formats around them.  The reference's own
versions of these two subroutines need MPI (distributeaveu/distributebondu,
communication_helper_real.f95), so they are pinned here by what the
mathematics fixes:

* aveflow on fields with separable time series u = a_t F(x), w = b_t G(x),
  v = c_t H(x), p = d_t P(x), sm = e_t S(x) must give means mean(a) F,
  population deviations |F| std(a), sm^2 means mean(e^2) S^2 and the flux
  uwfx = 0.25 (F(k-1)+F(k)) (G(i,k-1)+G(i+1,k-1)) cov(a, b): an index, a
  factor or a term dropped or moved fails one of them;
* averaging starts at n1 (earlier steps and earlier contents are
  discarded), the arrays' cells outside aveflow's loops are left alone, and
  avel ends at zero;
* bondv1_data24 sets u,v,w(i=0,1) to the planes, copies them over the
  domain at n == 1 and otherwise behaves as bondv1 (which is pinned against
  the reference build elsewhere).
"""
import numpy as np
import pytest

from lesamd import fortran_io, layout, setup_gen
from oracle import oracle

IM, JM, KM = 5, 4, 3


def _series(rng, nsteps):
    return {k: rng.uniform(0.5, 2.0, nsteps) for k in "abcde"}


def _spatial(rng, dtype):
    sp = {}
    for n in ("u", "v", "w", "p", "sm"):
        a = layout.new_field(n, IM, JM, KM, np.float64)
        a[...] = rng.uniform(-1.0, 1.0, a.shape)
        sp[n] = a
    return sp


def _run(prec, n1, nmax, nsteps, rng, fill=0.0):
    dt = np.float32 if prec == 32 else np.float64
    ts = _series(rng, nsteps)
    sp = _spatial(rng, dt)
    stats = {n: np.asfortranarray(np.full(layout.new_field(n, IM, JM, KM).shape, fill, dt))
             for n in layout.AVE_FIELDS}
    key = dict(u="a", w="b", v="c", p="d", sm="e")
    for n in range(1, nsteps + 1):
        fields = {f: np.asfortranarray((ts[key[f]][n - 1] * sp[f]).astype(dt)) for f in sp}
        oracle.oracle_aveflow(n, n1, nmax, stats, fields, IM, JM, KM, prec=prec)
    return ts, sp, stats


def _mean(x, n1, nmax):
    return x[n1 - 1:nmax].mean()


def _closed_form(ts, sp, n1, nmax):
    m = {k: _mean(v, n1, nmax) for k, v in ts.items()}
    var = {k: _mean((v - m[k]) ** 2, n1, nmax) for k, v in ts.items()}  # population variance
    cov_ab = _mean((ts["a"] - m["a"]) * (ts["b"] - m["b"]), n1, nmax)
    F, G, H, P, S = (sp[n] for n in ("u", "w", "v", "p", "sm"))
    # reference-layout index helpers: u,v (0:,-1:,0:), w (0:,-1:,-1:), p (0:,0:,0:), sm (-1:,-1:,0:)
    i = np.arange(1, IM + 1)[:, None, None]
    j = np.arange(1, JM + 1)[None, :, None]
    k = np.arange(1, KM + 1)[None, None, :]
    k0 = np.arange(0, KM + 1)[None, None, :]
    Fu = lambda ii, jj, kk: F[ii, jj + 1, kk]          # noqa: E731
    Gw = lambda ii, jj, kk: G[ii, jj + 1, kk + 1]      # noqa: E731
    out = {
        "aveu": m["a"] * Fu(i, j, k0),
        "avew": m["b"] * Gw(np.arange(1, IM + 2)[:, None, None], j, k0),
        "avev": m["c"] * H[i, j + 1, k],
        "avep": m["d"] * P[i, j, k],
        "avesm": m["e"] * S[i + 1, j + 1, k],
        "avesmsm": _mean(ts["e"] ** 2, n1, nmax) * S[i + 1, j + 1, k] ** 2,
        "aveuu": np.sqrt(var["a"]) * np.abs(Fu(i, j, k)),
        "avevv": np.sqrt(var["c"]) * np.abs(H[i, j + 1, k]),
        "aveww": np.sqrt(var["b"]) * np.abs(Gw(i, j, k)),
        "uwfx": 0.25 * (Fu(i, j, k - 1) + Fu(i, j, k)) * (Gw(i, j, k - 1) + Gw(i + 1, j, k - 1)) * cov_ab,
        "avel": np.zeros((IM, JM, KM)),
    }
    return out


def _written(name, a):
    """The cells aveflow's loops cover, from a statistics array in reference layout."""
    if name in ("aveu",):
        return a[:, :, 0:KM + 1]
    if name == "avew":
        return a[:, :, 0:KM + 1]
    if name == "avev":
        return a[:, :, 1:KM + 1]
    return a


@pytest.mark.parametrize("n1,nmax,nsteps", [(1, 4, 4), (2, 5, 5), (3, 3, 3)])
def test_aveflow_closed_forms_fp64(n1, nmax, nsteps):
    rng = np.random.default_rng(7 + n1)
    ts, sp, stats = _run(64, n1, nmax, nsteps, rng, fill=1e6)
    want = _closed_form(ts, sp, n1, nmax)
    for n in layout.AVE_FIELDS:
        got = _written(n, stats[n])
        np.testing.assert_allclose(got, want[n], rtol=1e-9, atol=1e-12, err_msg=n)


def test_aveflow_leaves_other_cells_alone():
    rng = np.random.default_rng(3)
    _, _, stats = _run(64, 2, 4, 4, rng, fill=-7.0)
    assert np.all(stats["avev"][:, :, 0] == -7.0)          # avev(k=0) is outside every loop
    assert np.all(stats["avew"][:, :, KM + 1:] == -7.0)    # avew(k=km+1..km+2) likewise


def test_aveflow_fp32_close_to_fp64():
    a = _run(32, 1, 4, 4, np.random.default_rng(11))[2]
    b = _run(64, 1, 4, 4, np.random.default_rng(11))[2]
    for n in layout.AVE_FIELDS:
        np.testing.assert_allclose(a[n], b[n], rtol=2e-5, atol=2e-6, err_msg=n)


def test_aveflow_before_n1_does_nothing():
    rng = np.random.default_rng(5)
    dt = np.float64
    stats = {n: np.asfortranarray(np.full(layout.new_field(n, IM, JM, KM).shape, 3.0, dt))
             for n in layout.AVE_FIELDS}
    fields = {f: np.asfortranarray(rng.uniform(-1, 1, layout.new_field(f, IM, JM, KM).shape))
              for f in ("u", "v", "w", "p", "sm")}
    oracle.oracle_aveflow(1, 2, 5, stats, fields, IM, JM, KM, prec=64)
    for n in layout.AVE_FIELDS:
        assert np.all(stats[n] == 3.0), n


def _small_setup():
    return setup_gen.build(setup_gen.small_workload(ustar=None, im=8, jm=6, km=5, seed=4))


@pytest.mark.parametrize("n", [1, 2])
def test_bondv1_data24_planes(n):
    s = _small_setup()
    w = s.workload
    rng = np.random.default_rng(n)
    ua, va, wa = (rng.uniform(-1, 3, (w.jm, w.km)) for _ in range(3))
    o = oracle.OracleBackend(s, prec=64)
    for f in ("u", "v", "w"):
        o.arrays[f][...] = rng.uniform(-1, 1, o.arrays[f].shape)
    before = {f: o.arrays[f].copy() for f in ("u", "v", "w")}
    o.bondv1_data24(n, ua, va, wa)
    u, v, wv = o.arrays["u"], o.arrays["v"], o.arrays["w"]
    # u, v (0:im+1,-1:jm+1,0:km+1), w (0:im+1,-1:jm+1,-1:km+1)
    last = w.im if n == 1 else 1
    for i in range(0, last + 1):
        if i == w.im:
            continue  # the outflow update rewrites u(im)
        np.testing.assert_array_equal(u[i, 2:w.jm + 2, 1:w.km + 1], ua, err_msg=f"u i={i}")
    for i in range(0, min(last, w.im) + 1):
        np.testing.assert_array_equal(v[i, 2:w.jm + 2, 1:w.km + 1], va, err_msg=f"v i={i}")
        np.testing.assert_array_equal(wv[i, 2:w.jm + 2, 2:w.km + 1], wa[:, :w.km - 1], err_msg=f"w i={i}")
        assert np.all(wv[i, :, w.km + 1] == 0.0)  # w(km) = 0 (bondv1.f95:190-195)
    if n != 1:  # interior columns 2..im-1 untouched
        np.testing.assert_array_equal(u[2:w.im, 2:w.jm + 2, 1:w.km + 1], before["u"][2:w.im, 2:w.jm + 2, 1:w.km + 1])


def test_bondv1_data24_with_u_plane_is_bondv1():
    s = _small_setup()
    w = s.workload
    rng = np.random.default_rng(9)
    prof = rng.uniform(1, 4, (w.jm, w.km))
    a = oracle.OracleBackend(s, prec=32)
    b = oracle.OracleBackend(s, prec=32)
    init = {f: rng.uniform(-1, 1, a.arrays[f].shape).astype(np.float32) for f in ("u", "v", "w")}
    for f in init:
        a.arrays[f][...] = init[f]
        b.arrays[f][...] = init[f]
    a.bondv1_data24(2, prof, np.zeros_like(prof), np.zeros_like(prof))
    prof32 = np.asfortranarray(prof.astype(np.float32))
    b._set_inflow(prof32.ctypes.data, w.jm, w.km)
    b.call("bondv1", n=2)
    b._set_inflow(None, 0, 0)
    for f in ("u", "v", "w"):
        np.testing.assert_array_equal(a.arrays[f], b.arrays[f], err_msg=f)


def test_data24_records(tmp_path):
    km, jm = 3, 7
    rec = fortran_io.DATA24_RECORD
    # value = 1e4 * (record number) + (index within the record), record numbers from 1
    planes = [np.asfortranarray(10000.0 * (np.arange(km)[None, :] + 1 + c * km) + np.arange(1, rec + 1)[:, None])
              for c in range(3)]
    ub, wb, vb = planes
    p = tmp_path / fortran_io.data24_name(5)
    assert p.name == "data24_out048005.dat"
    fortran_io.write_data24(p, ub, vb, wb)
    raw = p.read_bytes()
    assert len(raw) == 3 * km * rec * 4  # direct access: no record markers
    assert np.frombuffer(raw[:4], ">f4")[0] == 10001.0
    ua, va, wa = fortran_io.read_data24(p, jm, km)
    for k in range(km):
        for j in range(1, jm + 1):
            assert ua[j - 1, k] == 10000.0 * (k + 1) + 400 + j          # ub(400+j, k)
            assert wa[j - 1, k] == 10000.0 * (km + k + 1) + 400 + j     # w records follow u
            assert va[j - 1, k] == 10000.0 * (2 * km + k + 1) + 400 + j
    with pytest.raises(ValueError):
        fortran_io.read_data24(p, 801, km)


def test_aveflow_files(tmp_path):
    rng = np.random.default_rng(2)
    stats = {n: np.asfortranarray(rng.uniform(-1, 1, layout.new_field(n, IM, JM, KM).shape).astype(np.float32))
             for n in layout.AVE_FIELDS}
    fortran_io.write_aveflow(tmp_path / "data10", tmp_path / "data11", stats, KM)
    recs10 = fortran_io.read_records(tmp_path / "data10")
    recs11 = fortran_io.read_records(tmp_path / "data11")
    assert len(recs10) == 3 and len(recs11) == 4
    assert all(len(r) == 4 * IM * JM * KM for r in recs10 + recs11)
    # first value of data10 is aveu(1,1,1), big-endian
    assert np.frombuffer(recs10[0][:4], ">f4")[0] == stats["aveu"][0, 0, 1]
    # second record of data10 is avew over i=1..im (not im+1)
    assert np.frombuffer(recs10[1][4 * (IM - 1):4 * IM], ">f4")[0] == stats["avew"][IM - 1, 0, 1]
    assert np.frombuffer(recs10[1][4 * IM:4 * IM + 4], ">f4")[0] == stats["avew"][0, 1, 1]
    d10 = fortran_io.read_data10(tmp_path / "data10", IM, JM, KM)
    d11 = fortran_io.read_data11(tmp_path / "data11", IM, JM, KM)
    for n in fortran_io.DATA10:
        np.testing.assert_array_equal(d10[n], fortran_io.ave_interior(n, stats[n], KM))
    for n in fortran_io.DATA11:
        np.testing.assert_array_equal(d11[n], fortran_io.ave_interior(n, stats[n], KM))
