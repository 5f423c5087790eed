"""Domain set-up: grid metrics, Poisson coefficients, seeded city, masks and
the inflow profile, all in the reference's array layouts and FP32 arithmetic.

The reference builds these once on the host before the time loop:
  * metrics: grid.f95:36-172 (dx1=dy1=4 m, a 1 m x1.1 stretched dz rule);
  * Poisson coefficients cn1..cn4: init.f95:82-107;
  * building masks from a height map zbm: feedbfm.f95:87-116 (zbm itself
    is read from GIS files at feedbfm.f95:52-77);
  * inflow: a fixed 5 m/s at bondv1.f95:33-54.
The workloads used here replace the height map with a seeded synthetic
city, the dz rule with "dz1, x growth per level, capped", and the inflow
with a log-law profile; every formula that turns those inputs into the
arrays the hot path reads is the reference's, evaluated in the same FP32
order.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import layout

f32 = np.float32


@dataclasses.dataclass(frozen=True)
class CityParams:
    block: int = 8            # building block width in cells
    street: int = 2           # street width in cells
    occupancy: float = 1.0    # probability that a block holds a building
    median_h: float = 16.0    # lognormal median height [m]
    sigma: float = 0.5        # lognormal sigma


@dataclasses.dataclass(frozen=True)
class Workload:
    """One domain: sizes, grid rule, city, inflow, time stepping, solver."""
    name: str
    im: int
    jm: int
    km: int
    dx: float = 4.0
    dy: float = 4.0
    dz_rule: str = "stretched"   # "stretched" or "reference" (grid.f95:114-135)
    dz1: float = 2.0
    dz_growth: float = 1.05
    dz_cap: float = 20.0
    city: CityParams | None = CityParams()
    ustar: float | None = 0.4    # None -> the reference's fixed 5 m/s inflow
    z0: float = 0.1
    noise: float = 0.01
    dt: float = 0.05
    nsteps: int = 20
    nmaxp: int = 50
    pjuge: float = 1e-4
    omega: float = 1.0
    seed: int = 0

    @property
    def cells(self) -> int:
        return self.im * self.jm * self.km


# Constants of set.f95:89-95
RO = f32(1.1763)
VN = f32(15.83) * f32(10.0) ** f32(-6.0)   # set.f95:90, evaluated in FP32 as written
ALPHA = f32(-10.0)
BETA = f32(-1.0)


def workloads():
    """The named workloads (sizes and parameters of the benchmark configs)."""
    return {
        "plain-150x150x90": Workload("plain-150x150x90", 150, 150, 90, city=CityParams(8, 2, 1.0, 16.0, 0.5),
                                     ustar=0.4, z0=0.1, dt=0.05, nsteps=20),
        "plain-300x300x90": Workload("plain-300x300x90", 300, 300, 90, city=CityParams(8, 2, 1.0, 16.0, 0.5),
                                     ustar=0.5, z0=0.1, dt=0.05, nsteps=100),
        "large-1024x1024x128": Workload("large-1024x1024x128", 1024, 1024, 128, dx=2.0, dy=2.0,
                                        city=CityParams(10, 3, 0.35 / (10 / 13) ** 2, 20.0, 0.6),
                                        ustar=0.6, z0=0.3, dt=0.02, nsteps=100, nmaxp=100, pjuge=1e-5),
        "precision-512x512x128": Workload("precision-512x512x128", 512, 512, 128, dx=2.0, dy=2.0,
                                          city=CityParams(10, 3, 0.35 / (10 / 13) ** 2, 20.0, 0.6),
                                          ustar=0.6, z0=0.3, dt=0.02, nsteps=200, nmaxp=100, pjuge=1e-6),
    }


# ------------------------------------------------------------------ metrics
def grid_metrics(w: Workload) -> dict:
    """1-D metrics in reference layout (grid.f95:36-172)."""
    im, jm, km = w.im, w.jm, w.km
    m = {n: layout.new_metric(n, im, jm, km) for n in ("dx1", "dxl", "dxs", "dy1", "dyl", "dys", "dzn", "dzs", "z2")}
    m["dx1"][:] = f32(w.dx)                       # grid.f95:36-38, i=-1..ip+1
    dxl = m["dxl"]                                # grid.f95:50-53
    dxl[0] = f32(0.0)
    for i in range(1, im + 1):
        dxl[i] = dxl[i - 1] + m["dx1"][i + 1]
    m["dy1"][:] = f32(w.dy)                       # grid.f95:67-69, j=0..jp+1
    dyl = m["dyl"]                                # grid.f95:81-84
    dyl[0] = f32(0.0)
    for j in range(1, jm + 1):
        dyl[j] = dyl[j - 1] + m["dy1"][j]
    dzn = m["dzn"]   # index k -> dzn[k+1]
    z2 = m["z2"]     # index k -> z2[k]
    D = lambda k: k + 1  # noqa: E731
    lev = [f32(0.0)] * max(km + 2, 60)           # dzn(0..max(kp+1, 59))
    if w.dz_rule == "reference":                  # grid.f95:114-135
        lev[0] = lev[1] = f32(1.0)
        for k in range(2, 16):
            lev[k] = lev[k - 1] * f32(1.1)
        for k in range(16, 45):
            lev[k] = f32(4.0)
        for k in range(45, 59):
            lev[k] = lev[k - 1] * f32(1.1)
        for k in range(59, len(lev)):
            lev[k] = f32(16.0)
    elif w.dz_rule == "stretched":
        lev[0] = lev[1] = f32(w.dz1)
        for k in range(2, len(lev)):
            lev[k] = min(lev[k - 1] * f32(w.dz_growth), f32(w.dz_cap))
    else:
        raise ValueError(f"unknown dz rule {w.dz_rule!r}")
    z2[0] = lev[0]
    z2[1] = lev[1]
    for k in range(0, km + 2):
        dzn[D(k)] = lev[k]
    # z2(k)=z2(k-1)+dzn(k), k=2..kp+2, before the ghost copies below
    # (grid.f95:133-135; dzn(kp+2) is still zero there)
    for k in range(2, km + 3):
        z2[k] = z2[k - 1] + (dzn[D(k)] if k <= km + 1 else f32(0.0))
    dzn[D(km + 1)] = dzn[D(km)]                   # grid.f95:147-152
    dzn[D(km + 2)] = dzn[D(km + 1)]
    dzn[D(0)] = dzn[D(1)]
    dzn[D(-1)] = dzn[D(0)]
    dzs = m["dzs"]
    for k in range(0, km + 1):                    # grid.f95:154-156
        dzs[D(k)] = dzn[D(k + 1)] / f32(2.0) + dzn[D(k)] / f32(2.0)
    for i in range(0, im + 1):                    # grid.f95:160-162
        m["dxs"][i] = m["dx1"][i + 1] / f32(2.0) + m["dx1"][i + 2] / f32(2.0)
    for j in range(0, jm + 1):                    # grid.f95:166-168
        m["dys"][j] = m["dy1"][j] / f32(2.0) + m["dy1"][j + 1] / f32(2.0)
    dzs[D(km + 1)] = dzs[D(km)]                   # grid.f95:170-172
    dzs[D(km + 2)] = dzs[D(km + 1)]
    dzs[D(-1)] = dzs[D(0)]
    return m


def poisson_coefficients(m: dict, im: int, jm: int, km: int, with_cn1: bool = True) -> dict:
    """cn1..cn4 (init.f95:82-107), FP32 in the reference's order."""
    dxs, dys, dzs = m["dxs"], m["dys"], m["dzs"]
    c = {n: layout.new_metric(n, im, jm, km) for n in ("cn2l", "cn2s", "cn3l", "cn3s", "cn4l", "cn4s")}
    two, one = f32(2.0), f32(1.0)
    for i in range(1, im + 1):
        c["cn2s"][i - 1] = two / (dxs[i - 1] * (dxs[i - 1] + dxs[i]))
        c["cn2l"][i - 1] = two / (dxs[i] * (dxs[i - 1] + dxs[i]))
    for j in range(1, jm + 1):
        c["cn3s"][j - 1] = two / (dys[j - 1] * (dys[j - 1] + dys[j]))
        c["cn3l"][j - 1] = two / (dys[j] * (dys[j - 1] + dys[j]))
    ax = np.array([two / (dxs[i - 1] * dxs[i]) for i in range(1, im + 1)], dtype=f32)
    by = np.array([two / (dys[j - 1] * dys[j]) for j in range(1, jm + 1)], dtype=f32)
    cz = np.zeros(km, dtype=f32)
    for k in range(1, km + 1):
        dz1 = dzs[k - 1 + 1]
        dz2 = dzs[k + 1]
        c["cn4s"][k - 1] = two / (dz1 * (dz1 + dz2))
        c["cn4l"][k - 1] = two / (dz2 * (dz1 + dz2))
        cz[k - 1] = two / (dz1 * dz2)
    # cn1 = 1/(2/(dxs dxs) + 2/(dys dys) + 2/(dz1 dz2)), summed left to right
    if with_cn1:
        s = np.empty((im, jm, km), dtype=f32, order="F")
        np.add(ax[:, None, None], by[None, :, None], out=s)
        s += cz[None, None, :]
        np.divide(one, s, out=s)
        c["cn1"] = s
    c["_cn1_terms"] = (ax, by, cz)
    return c


# --------------------------------------------------------------------- city
def city_heights(w: Workload) -> np.ndarray:
    """zbm(i,j) for i=1..im, j=1..jm: blocks of block x block cells separated
    by street-wide streets, each block a building of lognormal height with
    probability `occupancy` (seeded)."""
    zb = np.zeros((w.im, w.jm), dtype=f32)
    c = w.city
    if c is None:
        return zb
    rng = np.random.default_rng(w.seed)
    period = c.block + c.street
    nbi = -(-w.im // period)
    nbj = -(-w.jm // period)
    heights = np.exp(rng.normal(math.log(c.median_h), c.sigma, size=(nbi, nbj)))
    occupied = rng.random((nbi, nbj)) < c.occupancy
    heights = np.where(occupied, heights, 0.0).astype(f32)
    lead = c.street // 2
    q_i = np.arange(w.im) % period
    q_j = np.arange(w.jm) % period
    in_i = (q_i >= lead) & (q_i < lead + c.block)
    in_j = (q_j >= lead) & (q_j < lead + c.block)
    bi = np.arange(w.im) // period
    bj = np.arange(w.jm) // period
    zb = np.where(in_i[:, None] & in_j[None, :], heights[bi[:, None], bj[None, :]], f32(0.0))
    return zb.astype(f32)


def building_masks(zbm: np.ndarray, m: dict, im: int, jm: int, km: int) -> dict:
    """amask1 and b/c/dmask1 (feedbfm.f95:87-116): amask1=0 where
    zbm(i,j) > z2(k)+0.5*dzn(k) on the interior, b=c=d=1-amask1 there,
    every ghost cell zero."""
    out = {n: layout.new_field(n, im, jm, km) for n in ("amask1", "bmask1", "cmask1", "dmask1")}
    z2 = m["z2"]
    dzn = m["dzn"]
    top = np.array([z2[k] + f32(0.5) * dzn[k + 1] for k in range(1, km + 1)], dtype=f32)
    solid = np.asfortranarray(zbm[:, :, None] > top[None, None, :])    # (im, jm, km)
    b = solid.astype(f32)
    out["amask1"][layout.ref_slice("amask1", (1, 1, 1), (im, jm, km))] = f32(1.0) - b
    for n in ("bmask1", "cmask1", "dmask1"):
        out[n][layout.ref_slice(n, (1, 1, 1), (im, jm, km))] = b
    return out


def inflow_profile(w: Workload, m: dict) -> np.ndarray | None:
    """u_in(j,k) = (u*/0.4) ln(z_c/z0) (1 + U[-noise, noise]) with
    z_c(k) = z2(k) + 0.5 dzn(k) (the centre height used at bondv1.f95:36);
    drawn once per run, j fastest.  None selects the fixed 5 m/s inflow."""
    if w.ustar is None:
        return None
    z2, dzn = m["z2"], m["dzn"]
    zc = np.array([float(z2[k]) + 0.5 * float(dzn[k + 1]) for k in range(1, w.km + 1)])
    rng = np.random.default_rng(w.seed + 7919)
    noise = rng.uniform(-w.noise, w.noise, size=(w.jm, w.km))
    prof = (w.ustar / 0.4) * np.log(zc / w.z0)[None, :] * (1.0 + noise)
    return np.asfortranarray(prof.astype(f32))


# ------------------------------------------------------------- whole domain
@dataclasses.dataclass
class DomainSetup:
    workload: Workload
    metrics: dict
    coeffs: dict
    masks: dict
    zbm: np.ndarray
    inflow: np.ndarray | None

    def new_state(self, dtype=np.float32) -> dict:
        """Zero-initialised per-step arrays (init.f95:46-69 zeroes u,v,w,p;
        every other array is zero static storage in main.f95)."""
        w = self.workload
        return {n: layout.new_field(n, w.im, w.jm, w.km, dtype) for n in layout.STEP_FIELDS}


def build(w: Workload, dense: bool = True) -> DomainSetup:
    """The set-up arrays of a workload.  dense=False leaves out the 3-D
    arrays (cn1 and the masks): the library then evaluates cn1 from the 1-D
    coefficients and builds the masks on the device from zbm."""
    m = grid_metrics(w)
    c = poisson_coefficients(m, w.im, w.jm, w.km, with_cn1=dense)
    zb = city_heights(w)
    masks = building_masks(zb, m, w.im, w.jm, w.km) if dense else None
    prof = inflow_profile(w, m)
    return DomainSetup(w, m, c, masks, zb, prof)


def small_workload(im=24, jm=20, km=16, seed=0, **kw) -> Workload:
    """A small domain with the plain-workload generator (test sizes)."""
    base = dict(city=CityParams(4, 2, 0.8, 6.0, 0.5), ustar=0.4, z0=0.1, dt=0.05, nsteps=3, seed=seed)
    base.update(kw)
    return Workload(f"small-{im}x{jm}x{km}", im, jm, km, **base)
