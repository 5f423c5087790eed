"""Pins of the restatement by mathematics rather than by the reference:

* SOR converges to the solution of the discrete Poisson problem built
  independently here as a matrix (Neumann in i and k, periodic in j, as the
  ghost copies of boundp1/boundp2 make it), solved by least squares;
* uniform flow U0 gives zero velocity gradients everywhere except the
  wall layer, where vel2's wall function is the log-law shear
  U0 / (z ln(z/z0)) at z = dzn(1)/2, and les's eddy viscosity there follows
  from it (including the zero diu3(0,.,.) the reference reads at i=1);
* the pressure returned by press has zero volume-weighted mean.
"""
import numpy as np

from lesamd import layout, setup_gen
from oracle import oracle


def _uniform_domain(im, jm, km, **kw):
    base = dict(dx=4.0, dy=4.0, dz1=2.0, dz_growth=1.0, city=None, ustar=None, dt=0.05)
    base.update(kw)
    return setup_gen.build(setup_gen.Workload("m", im, jm, km, **base))


def _poisson_matrix(s):
    """(1/cn1) p - sum_nb c_nb p_nb, ghosts mapped as boundp1/boundp2 map them."""
    w = s.workload
    im, jm, km = w.im, w.jm, w.km
    c = s.coeffs
    n = im * jm * km
    M = np.zeros((n, n))
    idx = lambda i, j, k: (i - 1) + (j - 1) * im + (k - 1) * im * jm  # noqa: E731
    for k in range(1, km + 1):
        for j in range(1, jm + 1):
            for i in range(1, im + 1):
                r = idx(i, j, k)
                M[r, r] += 1.0 / float(c["cn1"][i - 1, j - 1, k - 1])
                nbs = [
                    (float(c["cn2l"][i - 1]), (min(i + 1, im), j, k)),
                    (float(c["cn2s"][i - 1]), (max(i - 1, 1), j, k)),
                    (float(c["cn3l"][j - 1]), (i, j + 1 if j < jm else 1, k)),
                    (float(c["cn3s"][j - 1]), (i, j - 1 if j > 1 else jm, k)),
                    (float(c["cn4l"][k - 1]), (i, j, min(k + 1, km))),
                    (float(c["cn4s"][k - 1]), (i, j, max(k - 1, 1))),
                ]
                for coef, (a, b, d) in nbs:
                    M[r, idx(a, b, d)] -= coef
    return M


def test_sor_converges_to_direct_solution():
    s = _uniform_domain(8, 6, 5, nmaxp=3000, pjuge=1e-30)
    w = s.workload
    orc = oracle.OracleBackend(s, 64, redmode=1)
    rng = np.random.default_rng(0)
    r = rng.standard_normal((w.im, w.jm, w.km))
    # rhs = div(u)/dt with v=w=f=g=h=0 (press.f95:49-60): u(i)-u(i-1) = r dt dx1(i)
    u = orc.arrays["u"]
    dx1 = s.metrics["dx1"][2: w.im + 2].astype(np.float64)
    col = lambda i: layout.ref_slice("u", (i, 1, 1), (i, w.jm, w.km))  # noqa: E731
    for i in range(1, w.im + 1):
        u[col(i)] = u[col(i - 1)] + r[i - 1][None] * w.dt * dx1[i - 1]
    orc.call("press", n=1, nmax=1)
    st = orc.press_stats()
    assert st["iters"] == 3000
    p = layout.interior("p", orc.fetch(["p"])["p"], w.im, w.jm, w.km)
    rhs = r - r.mean()  # uniform grid: the volume-weighted mean is the plain mean
    M = _poisson_matrix(s)
    x = np.linalg.lstsq(M, -rhs.reshape(-1, order="F"), rcond=None)[0]
    x = x - x.mean()
    np.testing.assert_allclose(p.reshape(-1, order="F"), x, rtol=0, atol=1e-9 * np.abs(x).max())
    # residual history decreases
    h = st["sor"]
    assert h[-1] < 1e-15 * h[0]


def test_uniform_flow_wall_layer_closed_form():
    s = _uniform_domain(7, 5, 6)
    w = s.workload
    orc = oracle.OracleBackend(s, 64, redmode=1)
    U0 = 5.0
    orc.arrays["u"][...] = U0
    orc.call("velfg")
    orc.call("les")
    f = orc.fetch(["uspd", "sm"] + [f"diu{q}" for q in range(1, 10)])
    im, jm, km = w.im, w.jm, w.km
    dz1 = float(s.metrics["dzn"][2])
    z = 0.5 * dz1
    shear = U0 / (z * np.log(z / 0.1))      # log-law: u*/(kappa z), u* = kappa U0 / ln(z/z0)
    d3 = layout.interior("diu3", f["diu3"], im, jm, km)
    np.testing.assert_allclose(d3[:, :, 0], shear, rtol=1e-12)
    assert np.all(d3[:, :, 1:] == 0)
    for q in (1, 2, 4, 5, 6, 7, 8, 9):
        assert not layout.interior(f"diu{q}", f[f"diu{q}"], im, jm, km).any(), q
    np.testing.assert_array_equal(layout.interior("uspd", f["uspd"][:, :, None], im, jm, 1)[:, :, 0], 1.0)
    sm = layout.interior("sm", f["sm"], im, jm, km)
    delta = (4.0 * 4.0 * dz1) ** (1.0 / 3.0)
    # k=1: dudz = (diu3(i-1,j,1)+diu3(i,j,1))/4 with diu3(0,j,1)=0
    np.testing.assert_allclose(sm[1:, :, 0], (0.1 * delta) ** 2 * shear / 2, rtol=1e-12)
    np.testing.assert_allclose(sm[0, :, 0], (0.1 * delta) ** 2 * shear / 4, rtol=1e-12)
    assert np.all(sm[:, :, 1:] == 0)


def test_press_removes_volume_weighted_mean():
    s = setup_gen.build(setup_gen.small_workload(12, 10, 9, seed=3, ustar=None))
    w = s.workload
    orc = oracle.OracleBackend(s, 64, redmode=1)
    for n in (1, 2):
        orc.step(n, 2)
    p = layout.interior("p", orc.fetch(["p"])["p"], w.im, w.jm, w.km)
    dzn = s.metrics["dzn"][2: w.km + 2].astype(np.float64)
    dv = 16.0 * dzn[None, None, :]
    assert abs((p * dv).sum()) <= 1e-12 * (np.abs(p) * dv).sum()
