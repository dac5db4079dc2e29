"""Seeded synthetic domains: grid, D8 river network, land cover, soils,
daily forcing and parameter sets, in the reference's array conventions.

Grid cells are numbered column-major (row index fastest) and every cell of
the rectangle is active, so L1 cell k is L11 node k (L0 = L1 = L11).  Links
follow L11_set_network_topology (src/mRM/mo_mrm_net_startup.f90:630-698):
one link per non-outlet node, in node order; netPerm follows
L11_routing_order (728-859).  Link length is the cell size (diagonal: times
sqrt 2) and slope the drop over it, floored at 1e-4 as at 1428-1429.

Forcing (daily): wet days ~ Bernoulli(0.4) with Gamma(k=0.6, theta=4 mm/d)
precipitation; tavg = 10 + 12 sin(2 pi (doy - 105)/365) + N(0, 3) degC;
tmax/tmin = tavg +- 5 degC.  The reference's default monthly night/day
ratios and pan-evaporation coefficients (mhm.nml:788-813) are used.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import parameters as PM

# mhm.nml:788 (panEvapo) and 794-813 (nightDayRatio)
EVAP_COEFF = np.array([1.30, 1.20, 0.72, 0.75, 1.00, 1.00, 1.00, 1.00, 1.00, 1.00, 1.00, 1.50])
FNIGHT_PREC = np.array([0.46, 0.50, 0.52, 0.51, 0.48, 0.50, 0.49, 0.48, 0.52, 0.56, 0.50, 0.47])
FNIGHT_PET = np.full(12, 0.10)
FNIGHT_TEMP = np.array([-0.76, -1.30, -1.88, -2.38, -2.72, -2.75, -2.74, -3.04, -2.44, -1.60,
                        -0.94, -0.53])
HORIZON_DEPTH_3 = np.array([200.0, 500.0, 1000.0])


@dataclass
class Domain:
    nrows: int
    ncols: int
    cellsize: float
    seed: int
    horizon_depth: np.ndarray
    lat: np.ndarray
    fractions: np.ndarray
    texture: np.ndarray
    aspect: np.ndarray
    elev: np.ndarray
    fdir: np.ndarray
    fromN: np.ndarray
    toN: np.ndarray
    netPerm: np.ndarray
    rOrder: np.ndarray
    length: np.ndarray
    slope: np.ndarray
    fFPimp: np.ndarray
    area_km2: np.ndarray
    depth: np.ndarray
    gauge_nodes: np.ndarray
    n_outlets: int = 1
    forcing: dict = field(default_factory=dict)
    n_days: int = 0
    # routing on a coarser L11 grid: the L11 node (1-based) of every L1 cell
    # (L1_L11_Id) and the node count; None: L1 = L11
    l1_l11: np.ndarray = None
    nodes_l11: int = 0

    @property
    def n_cells(self):
        return self.nrows * self.ncols

    @property
    def n_nodes(self):
        return self.nodes_l11 if self.l1_l11 is not None else self.n_cells

    @property
    def n_links(self):
        return self.fromN.shape[0]

    @property
    def dmax(self):
        return int(self.depth.max())


def _fbm(rng, nrows, ncols, hurst):
    ky = np.fft.fftfreq(nrows)[:, None]
    kx = np.fft.rfftfreq(ncols)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    spec = amp * (rng.normal(size=amp.shape) + 1j * rng.normal(size=amp.shape))
    z = np.fft.irfft2(spec, s=(nrows, ncols))
    z -= z.min()
    return z / max(z.max(), 1e-12)


def make_domain(nrows, ncols, seed, cellsize=1000.0, horizon_depth=HORIZON_DEPTH_3, hurst=0.8,
                relief=500.0, lat0=48.0):
    """Terrain, network, land cover and soils of one basin (no forcing)."""
    rng = np.random.default_rng(seed)
    z = relief * _fbm(rng, nrows, ncols, hurst)
    # gentle tilt towards the bottom-centre outlet so the main stem spans the grid
    ii, jj = np.meshgrid(np.arange(nrows), np.arange(ncols), indexing="ij")
    z = z + 0.2 * relief * (nrows - 1 - ii) / max(nrows - 1, 1)
    z = z + 0.05 * relief * np.abs(jj - ncols // 2) / max(ncols, 1)
    dem = np.asfortranarray(z).ravel(order="F").astype(np.float64)
    outlet = (nrows - 1) + (ncols // 2) * nrows
    dem[outlet] = z.min() - 1.0
    n = nrows * ncols
    filled = np.zeros(n)
    down = np.zeros(n, np.int64)
    fdir = np.zeros(n, np.int32)
    rc = N.syn().syn_priority_flood(nrows, ncols, N.dptr(dem), outlet, 1e-3, N.dptr(filled),
                                    N.lptr(down), N.iptr(fdir))
    assert rc == 0
    depth = np.zeros(n, np.int32)
    assert N.syn().syn_tree_depth(n, N.lptr(down), N.iptr(depth)) == 0
    src = np.nonzero(down >= 0)[0]
    fromN = (src + 1).astype(np.int32)
    toN = (down[src] + 1).astype(np.int32)
    r_order, net_perm = N.routing_order(fromN, toN, n)
    n_links = src.shape[0]
    diag = np.isin(fdir[src], [2, 8, 32, 128])
    length = np.where(diag, cellsize * np.sqrt(2.0), cellsize)
    slope = (filled[src] - filled[down[src]]) / length
    slope = np.where(slope < 1.0e-4, 1.0e-4, slope)
    # land cover: each L1 cell aggregates 2 x 2 L0 cells of 3 uniform classes
    cls = rng.integers(0, 3, size=(n, 4))
    fractions = np.stack([(cls == c).mean(axis=1) for c in range(3)])
    texture = np.clip(_fbm(rng, nrows, ncols, 0.5).ravel(order="F"), 0.0, 0.999999)
    if nrows >= 2 and ncols >= 2:
        gy, gx = np.gradient(z, cellsize)
        aspect = (np.degrees(np.arctan2(-gx, gy)) % 360.0).ravel(order="F")
    else:
        aspect = np.full(n, 180.0)
    lat = lat0 + (nrows / 2.0 - ii.ravel(order="F")) * cellsize / 111000.0
    area = np.full(n, cellsize * cellsize * 1.0e-6)
    return Domain(nrows=nrows, ncols=ncols, cellsize=cellsize, seed=seed,
                  horizon_depth=np.asarray(horizon_depth, dtype=np.float64), lat=lat,
                  fractions=fractions, texture=texture, aspect=aspect, elev=filled, fdir=fdir,
                  fromN=fromN, toN=toN, netPerm=net_perm[:n_links].copy(),
                  rOrder=r_order[:n_links].copy(), length=length, slope=slope,
                  fFPimp=fractions[1][src].copy(), area_km2=area, depth=depth,
                  gauge_nodes=np.array([outlet + 1], np.int32))


def refine_domain(dom11, f, seed):
    """A basin whose hydrology (L1) grid is f x f times finer than the routing
    grid of dom11 (L11, resolution_Hydrology = resolution_Routing / f as in
    the reference's check case_04): the network, gauges and links of dom11,
    L1 land cover, soils, terrain aspect and latitude on the fine grid, and
    L1_L11_Id mapping fine cell (i, j) to coarse node (i // f, j // f)
    (column-major numbering, 1-based)."""
    rng = np.random.default_rng(seed)
    nr, nc = dom11.nrows * f, dom11.ncols * f
    n = nr * nc
    ii, jj = np.meshgrid(np.arange(nr), np.arange(nc), indexing="ij")
    i = ii.ravel(order="F")
    j = jj.ravel(order="F")
    l1_l11 = ((j // f) * dom11.nrows + (i // f) + 1).astype(np.int32)
    cls = rng.integers(0, 3, size=(n, 4))
    fractions = np.stack([(cls == c).mean(axis=1) for c in range(3)])
    texture = np.clip(_fbm(rng, nr, nc, 0.5).ravel(order="F"), 0.0, 0.999999)
    aspect = rng.uniform(0.0, 360.0, n)
    cs = dom11.cellsize / f
    lat = dom11.lat[l1_l11 - 1] + ((i % f) - (f - 1) / 2.0) * cs / 111000.0
    return Domain(nrows=nr, ncols=nc, cellsize=cs, seed=seed, horizon_depth=dom11.horizon_depth,
                  lat=lat, fractions=fractions, texture=texture, aspect=aspect,
                  elev=dom11.elev[l1_l11 - 1], fdir=dom11.fdir, fromN=dom11.fromN, toN=dom11.toN,
                  netPerm=dom11.netPerm, rOrder=dom11.rOrder, length=dom11.length,
                  slope=dom11.slope, fFPimp=dom11.fFPimp, area_km2=np.full(n, cs * cs * 1.0e-6),
                  depth=dom11.depth, gauge_nodes=dom11.gauge_nodes, n_outlets=dom11.n_outlets,
                  l1_l11=l1_l11, nodes_l11=dom11.n_cells)


def make_forcing(dom, n_days, seed, doy0=1):
    """Daily pre/tavg/tmin/tmax, each (n_days, n_cells) = Fortran L1_x(n_cells, n_days)."""
    rng = np.random.default_rng(seed)
    n = dom.n_cells
    wet = rng.random((n_days, n)) < 0.4
    pre = np.where(wet, rng.gamma(0.6, 4.0, size=(n_days, n)), 0.0)
    doy = (doy0 - 1 + np.arange(n_days)) % 365 + 1
    tavg = 10.0 + 12.0 * np.sin(2.0 * np.pi * (doy - 105) / 365.0)[:, None] + rng.normal(0.0, 3.0, (n_days, n))
    dom.forcing = {"PRE": np.ascontiguousarray(pre), "TAVG": np.ascontiguousarray(tavg),
                   "TMIN": np.ascontiguousarray(tavg - 5.0),
                   "TMAX": np.ascontiguousarray(tavg + 5.0)}
    dom.n_days = n_days
    return dom.forcing


def cell_params(dom, gamma, pet_case=1):
    """L1 effective parameters of one gamma vector (MPR-lite) plus latitude;
    pet_case 2 or 3 adds that case's monthly PET parameters."""
    st = getattr(dom, "_mpr_static", None)
    if st is None:
        st = dom._mpr_static = PM.static_fields(dom.fractions, dom.texture, dom.horizon_depth)
    p = PM.mpr_lite(gamma, dom.fractions, dom.texture, dom.aspect, dom.horizon_depth, static=st)
    if pet_case >= 2:
        p.update(PM.pet_radiation_params(st["LAI"], pet_case))
    p["LATITUDE"] = dom.lat
    return p


def make_pet_forcing(dom, seed, doy0=1):
    """The radiation-based PET cases' daily inputs beside make_forcing's
    (same days): net radiation [W m-2] = 70 + 90 sin(2 pi (doy - 80)/365) +
    N(0, 25), negative on some winter days; actual vapour pressure [Pa] =
    RH * 1000 * 0.6108 exp(17.27 tavg / (tavg + 237.3)) with RH ~ U(0.35, 0.95);
    wind speed [m s-1] ~ 0.3 + Gamma(k=2, theta=1.2)."""
    rng = np.random.default_rng(seed)
    n, nd = dom.n_cells, dom.n_days
    doy = (doy0 - 1 + np.arange(nd)) % 365 + 1
    rn = 70.0 + 90.0 * np.sin(2.0 * np.pi * (doy - 80) / 365.0)[:, None] + rng.normal(0.0, 25.0, (nd, n))
    tavg = dom.forcing["TAVG"]
    es = 0.6108 * np.exp(17.27 * tavg / (tavg + 237.3))
    avp = rng.uniform(0.35, 0.95, (nd, n)) * 1000.0 * es
    ws = 0.3 + rng.gamma(2.0, 1.2, (nd, n))
    dom.forcing.update(NETRAD=np.ascontiguousarray(rn), ABSVAPPRESS=np.ascontiguousarray(avp),
                       WINDSPEED=np.ascontiguousarray(ws))
    return dom.forcing


def make_meteo_weights(n_cells, seed):
    """Per-cell meteo weights (n_cells, 12, 24) for pre, temp and pet in the
    reference's sense (L1_x_weights(cell, month, hour + 1)): precipitation
    and PET weights sum to 1 over the day's hours (PET peaking at noon), the
    temperature weights scale the absolute temperature around 1."""
    rng = np.random.default_rng(seed)
    h = np.arange(24)
    pre = rng.gamma(1.0, 1.0, (n_cells, 12, 24))
    pre /= pre.sum(axis=2, keepdims=True)
    shape = np.maximum(np.sin(np.pi * (h - 5) / 15.0), 0.02)
    pet = shape[None, None, :] * rng.uniform(0.8, 1.2, (n_cells, 12, 24))
    pet /= pet.sum(axis=2, keepdims=True)
    temp = 1.0 + 0.012 * np.sin(2.0 * np.pi * (h - 9) / 24.0)[None, None, :] \
        + rng.normal(0.0, 0.002, (n_cells, 12, 24))
    return pre, temp, pet
