"""Seeded synthetic LETKF instances (inputs only; never part of the timed path).

One integer seed drives every draw through numpy's Philox-4x32-10 counter
RNG (Salmon et al. 2011), so an instance is bit-identical on every host.

Instance layout follows the reference arrays:
  * ``xb``  perturbations, Fortran ``xb(ens, n, lat, lon)``
    (x86/DA/LETKF/src/letkf_compute.F90:15) == C array ``[lon][lat][n][k]``;
  * ``axb`` mean ``axb(n, lat, lon)`` == C ``[lon][lat][n]``;
  * obs sites (lat, lon) each with a run of rows (level-major, then var, as
    letkf_yb.F90:102-103); ``yb`` rows are ``pro_yb(ens, .)`` centred as
    letkf_obs.F90:156-183: Yb = H(x_e) - mean_e H(x_e), dep = y - mean_e H(x_e);
  * ``rinv`` = 1/sigma^2 as read_obs_mod.F90:438.
The background is split into mean and perturbations as
module_da_parallel.F90:195-196.

Fields: truth = sum of 8 orthonormal real spherical harmonics (l <= 6,
N(0,1) coefficients); member = truth + N(0,1) noise smoothed by a Gaussian
kernel of the configured width (separable on the lat-lon grid: periodic FFT
filter along each latitude circle with the circle's metric, a normalised
Gaussian along meridians), rescaled to unit variance; 3-D variables carry a
vertical AR(1) with correlation 0.8 between adjacent levels.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .config import LetkfConfig, PREC_FP64, VCOORD_LEVEL

EARTH_KM = 6371.0


def _rng(seed: int, stream: int) -> np.random.Generator:
    ss = np.random.SeedSequence([int(seed) & 0xFFFFFFFF, int(stream)])
    return np.random.Generator(np.random.Philox(ss))


def cell_centred_grid(nlat: int, nlon: int):
    """Cell-centred lat-lon grid as real(4), like pro_info's grid_lat/grid_lon."""
    dlat = 180.0 / nlat
    dlon = 360.0 / nlon
    lat = (-90.0 + (np.arange(nlat) + 0.5) * dlat).astype(np.float32)
    lon = ((np.arange(nlon) + 0.5) * dlon).astype(np.float32)
    return lat, lon


def _real_sph_harm(l: int, m: int, lat_rad: np.ndarray, lon_rad: np.ndarray) -> np.ndarray:
    from scipy.special import sph_harm_y
    theta = (np.pi / 2 - lat_rad)[:, None]
    phi = lon_rad[None, :]
    am = abs(m)
    y = sph_harm_y(l, am, theta, phi)   # orthonormal complex Y_l^|m|
    if m == 0:
        return np.real(y)
    if m > 0:
        return math.sqrt(2.0) * (-1.0) ** m * np.real(y)
    return math.sqrt(2.0) * (-1.0) ** am * np.imag(y)


def truth_field(rng: np.random.Generator, lat: np.ndarray, lon: np.ndarray, nterms: int = 8,
                lmax: int = 6) -> np.ndarray:
    pairs = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
    pick = rng.choice(len(pairs), size=nterms, replace=False)
    coef = rng.standard_normal(nterms)
    la = np.deg2rad(lat.astype(np.float64))
    lo = np.deg2rad(lon.astype(np.float64))
    f = np.zeros((lat.size, lon.size))
    for c, p in zip(coef, pick):
        l, m = pairs[p]
        f += c * _real_sph_harm(l, m, la, lo)
    return f


class _Smoother:
    """Separable Gaussian smoother on a cell-centred lat-lon grid."""

    def __init__(self, lat: np.ndarray, lon: np.ndarray, sigma_km: float):
        nlat, nlon = lat.size, lon.size
        la = np.deg2rad(lat.astype(np.float64))
        dlon = 2 * np.pi / nlon
        m = np.arange(nlon // 2 + 1)
        coslat = np.maximum(np.cos(la), 1e-6)
        sig_cells = sigma_km / (EARTH_KM * coslat * dlon)
        self.tf = np.exp(-0.5 * (2 * np.pi * m[None, :] * sig_cells[:, None] / nlon) ** 2)
        dist = (la[:, None] - la[None, :]) * EARTH_KM
        K = np.exp(-0.5 * (dist / sigma_km) ** 2)
        self.K = K / K.sum(axis=1, keepdims=True)
        self.nlon = nlon

    def __call__(self, f: np.ndarray) -> np.ndarray:
        """f: (..., nlat, nlon) -> smoothed, unit variance per field."""
        F = np.fft.rfft(f, axis=-1)
        F *= self.tf
        g = np.fft.irfft(F, n=self.nlon, axis=-1)
        g = np.einsum('ij,...jk->...ik', self.K, g, optimize=True)
        sd = g.std(axis=(-2, -1), keepdims=True)
        return g / np.where(sd > 0, sd, 1.0)


@dataclass
class Obs:
    site_lat: np.ndarray            # f64 [n_sites]
    site_lon: np.ndarray            # f64 [n_sites]
    site_row_ptr: np.ndarray        # i32 [n_sites + 1]
    yb: np.ndarray                  # f64 [n_rows, k]
    dep: np.ndarray                 # f64 [n_rows]
    rinv: np.ndarray                # f64 [n_rows]
    vcoord: np.ndarray              # f64 [n_rows]
    row_var: np.ndarray             # i32 [n_rows] (state variable index)

    @property
    def n_sites(self) -> int:
        return int(self.site_lat.size)

    @property
    def n_rows(self) -> int:
        return int(self.dep.size)


@dataclass
class Instance:
    cfg: LetkfConfig
    grid_lat: np.ndarray            # f32 [nlat]
    grid_lon: np.ndarray            # f32 [nlon]
    xb: np.ndarray                  # f64 [nlon+2h, nlat+2h, n, k]
    axb: np.ndarray                 # f64 [nlon+2h, nlat+2h, n]
    truth: np.ndarray               # f64 [nlon, nlat, n]
    obs: Obs
    name: str = ""
    seed: int = 0
    meta: dict = field(default_factory=dict)


# --------------------------------------------------------------------------
# observation locations
# --------------------------------------------------------------------------

def _uniform_sphere(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    lon = rng.uniform(0.0, 360.0, n)
    return np.rad2deg(np.arcsin(z)), lon


def _swath_points(rng, n, n_swaths, width_km):
    """Points on n_swaths random great circles, uniform across a band of width_km."""
    plat, plon = _uniform_sphere(rng, n_swaths)
    which = rng.integers(0, n_swaths, n)
    t = rng.uniform(0.0, 2 * np.pi, n)
    off = rng.uniform(-0.5, 0.5, n) * width_km / EARTH_KM
    pl, po = np.deg2rad(plat[which]), np.deg2rad(plon[which])
    p = np.stack([np.cos(pl) * np.cos(po), np.cos(pl) * np.sin(po), np.sin(pl)], axis=1)
    ref = np.where(np.abs(p[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    e1 = np.cross(p, ref)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(p, e1)
    v = (np.cos(off)[:, None] * (np.cos(t)[:, None] * e1 + np.sin(t)[:, None] * e2)
         + np.sin(off)[:, None] * p)
    lat = np.rad2deg(np.arcsin(np.clip(v[:, 2], -1, 1)))
    lon = np.mod(np.rad2deg(np.arctan2(v[:, 1], v[:, 0])), 360.0)
    return lat, lon


# --------------------------------------------------------------------------
# observation operator (bilinear in lat/lon, linear in level index)
# --------------------------------------------------------------------------

def _h_weights(lat, lon, lev, var, cfg: LetkfConfig, grid_lat, grid_lon):
    """Corner state indices (interior [lon][lat][n] flat) and weights, [rows, 8]."""
    nlat, nlon = cfg.nlat, cfg.nlon
    dlat, dlon = 180.0 / nlat, 360.0 / nlon
    fy = (lat + 90.0) / dlat - 0.5
    i0 = np.floor(fy).astype(np.int64)
    wy = fy - i0
    i1 = i0 + 1
    i0c, i1c = np.clip(i0, 0, nlat - 1), np.clip(i1, 0, nlat - 1)
    fx = lon / dlon - 0.5
    j0 = np.floor(fx).astype(np.int64)
    wx = fx - j0
    j0m, j1m = np.mod(j0, nlon), np.mod(j0 + 1, nlon)
    var_start = cfg.var_start() - 1
    var_lev = cfg.var_lev()
    vs = var_start[var]
    is3d = var_lev[var] > 1
    l0 = np.where(is3d, np.clip(np.floor(lev).astype(np.int64), 1, cfg.nz), 1)
    l1 = np.where(is3d, np.clip(l0 + 1, 1, cfg.nz), 1)
    wz = np.where(is3d, np.clip(lev - l0, 0.0, 1.0), 0.0)
    n = cfg.n
    idx = np.empty((lat.size, 8), dtype=np.int64)
    w = np.empty((lat.size, 8))
    c = 0
    for (jj, wxx) in ((j0m, 1 - wx), (j1m, wx)):
        for (ii, wyy) in ((i0c, 1 - wy), (i1c, wy)):
            for (ll, wzz) in ((l0, 1 - wz), (l1, wz)):
                idx[:, c] = (jj * nlat + ii) * n + vs + (ll - 1)
                w[:, c] = wxx * wyy * wzz
                c += 1
    return idx, w


# --------------------------------------------------------------------------
# configurations
# --------------------------------------------------------------------------

PRESETS = {
    # c0: 72x36, one level and variable, k=20, 2,000 uniform obs (CPU-runnable)
    "c0": dict(nlat=36, nlon=72, nz=1, n2d=1, n3d=0, k=20, n_obs=2000, mix="uniform",
               sigma_obs="c0", radius_km=1000.0, smooth_km=500.0, vert=None),
    # c1: GMCORE-shaped 360x180x32, k=32, vertical localisation in level units
    "c1": dict(nlat=180, nlon=360, nz=32, n2d=1, n3d=3, k=32, n_obs=100_000, mix="uniform3d",
               sigma_obs="tuv", radius_km=800.0, smooth_km=500.0, vert=2.0),
    # c2: 360x180x32, clustered obs, k=64 (the benchmarked workload)
    "c2": dict(nlat=180, nlon=360, nz=32, n2d=1, n3d=3, k=64, n_obs=500_000, mix="clustered",
               sigma_obs="tuv", radius_km=600.0, smooth_km=500.0, vert=None),
    # c3: 720x360x32, k=64/128, 2M clustered obs, log-uniform sigma
    "c3": dict(nlat=360, nlon=720, nz=32, n2d=1, n3d=3, k=64, n_obs=2_000_000, mix="clustered",
               sigma_obs="loguniform", radius_km=600.0, smooth_km=500.0, vert=None),
}

# The configs quote a localisation "radius"; the reference parameter is
# da_loc_distance c with GC support 2c.  For the clustered configs the quoted
# local obs range (0 to ~3000 per column) is met with c = radius/2 (c = radius
# gives ~4,400 rows per column on average), so RADIUS_TO_C is 0.5 there.
RADIUS_TO_C = {"c0": 1.0, "c1": 1.0, "c2": 0.5, "c3": 0.5}


def make_config(name: str, k: Optional[int] = None, precision: int = PREC_FP64,
                **over) -> LetkfConfig:
    p = dict(PRESETS[name])
    p.update({kk: v for kk, v in over.items() if kk in p})
    cfg = LetkfConfig(
        k=k or p["k"], nlat=p["nlat"], nlon=p["nlon"], nz=p["nz"], n2d=p["n2d"], n3d=p["n3d"],
        local_dist=p["radius_km"] * RADIUS_TO_C.get(name, 1.0),
        vert_dist=p["vert"] if p["vert"] else -1.0,
        gc_filter=True, gc_filter_vertical=True, cube_n=1, cube_nz=1,
        precision=precision, vcoord_mode=VCOORD_LEVEL,
    )
    for kk, v in over.items():
        if hasattr(cfg, kk) and kk not in p:
            setattr(cfg, kk, v)
    return cfg


def generate(name: str, seed: int = 1, k: Optional[int] = None, precision: int = PREC_FP64,
             n_obs: Optional[int] = None, nlat: Optional[int] = None, nlon: Optional[int] = None,
             nz: Optional[int] = None, halo: int = 0, local_dist: Optional[float] = None,
             dtype_check: bool = True) -> Instance:
    """Generate config `name` ("c0".."c3"), optionally at a reduced size."""
    p = dict(PRESETS[name])
    if nlat: p["nlat"] = nlat
    if nlon: p["nlon"] = nlon
    if nz: p["nz"] = nz
    if n_obs: p["n_obs"] = n_obs
    cfg = make_config(name, k=k, precision=precision, nlat=p["nlat"], nlon=p["nlon"], nz=p["nz"])
    cfg.halo = halo
    if local_dist is not None:
        cfg.local_dist = float(local_dist)
    kk = cfg.k
    grid_lat, grid_lon = cell_centred_grid(cfg.nlat, cfg.nlon)
    nlat_, nlon_, n = cfg.nlat, cfg.nlon, cfg.n

    # ---- truth and members ----------------------------------------------
    rng_t = _rng(seed, 1)
    truth = np.empty((nlon_, nlat_, n))
    sm = _Smoother(grid_lat, grid_lon, p["smooth_km"])
    var_start = cfg.var_start() - 1
    var_lev = cfg.var_lev()
    for v in range(cfg.n_var):
        nl = int(var_lev[v])
        prev = None
        for l in range(nl):
            g = truth_field(rng_t, grid_lat, grid_lon)
            f = g if prev is None else 0.8 * prev + 0.6 * g
            truth[:, :, var_start[v] + l] = f.T
            prev = f
    members = np.empty((nlon_, nlat_, n, kk))
    rng_m = _rng(seed, 2)
    for v in range(cfg.n_var):
        nl = int(var_lev[v])
        prev = None
        for l in range(nl):
            g = sm(rng_m.standard_normal((kk, nlat_, nlon_)))
            f = g if prev is None else 0.8 * prev + 0.6 * g
            members[:, :, var_start[v] + l, :] = np.transpose(f, (2, 1, 0))
            prev = f
            del g
    members += truth[:, :, :, None]

    # ---- observation sites and rows -------------------------------------
    rng_o = _rng(seed, 3)
    nobs = int(p["n_obs"])
    mix = p["mix"]
    three_d = [v for v in range(cfg.n_var) if var_lev[v] > 1]
    if mix == "uniform":
        la, lo = _uniform_sphere(rng_o, nobs)
        rows_per = np.ones(nobs, dtype=np.int64)
        r_var = np.zeros(nobs, dtype=np.int64)
        r_lev = np.ones(nobs)
    elif mix == "uniform3d":
        la, lo = _uniform_sphere(rng_o, nobs)
        rows_per = np.ones(nobs, dtype=np.int64)
        r_var = np.asarray(three_d)[rng_o.integers(0, len(three_d), nobs)]
        r_lev = rng_o.uniform(1.0, float(cfg.nz), nobs)
    else:  # clustered: 70% swaths, 20% station columns, 10% uniform
        n_st = 1000 if nobs >= 100_000 else max(4, nobs // 500)
        per_st = cfg.nz * len(three_d)
        n_station_rows = n_st * per_st
        n_sw = int(round(0.7 * nobs))
        n_un = max(nobs - n_sw - n_station_rows, 0)
        sla, slo = _swath_points(rng_o, n_sw, 40, 50.0)
        tla, tlo = _uniform_sphere(rng_o, n_st)
        ula, ulo = _uniform_sphere(rng_o, n_un)
        la = np.concatenate([sla, tla, ula])
        lo = np.concatenate([slo, tlo, ulo])
        rows_per = np.concatenate([np.ones(n_sw, np.int64), np.full(n_st, per_st, np.int64),
                                   np.ones(n_un, np.int64)])
        sw_var = np.asarray(three_d)[rng_o.integers(0, len(three_d), n_sw)]
        sw_lev = rng_o.uniform(1.0, float(cfg.nz), n_sw)
        st_lev = np.tile(np.repeat(np.arange(1, cfg.nz + 1, dtype=np.float64), len(three_d)), n_st)
        st_var = np.tile(np.tile(np.asarray(three_d), cfg.nz), n_st)
        un_var = np.asarray(three_d)[rng_o.integers(0, len(three_d), n_un)]
        un_lev = rng_o.uniform(1.0, float(cfg.nz), n_un)
        r_var = np.concatenate([sw_var, st_var, un_var])
        r_lev = np.concatenate([sw_lev, st_lev, un_lev])
    n_sites = la.size
    row_ptr = np.zeros(n_sites + 1, dtype=np.int64)
    np.cumsum(rows_per, out=row_ptr[1:])
    n_rows = int(row_ptr[-1])
    row_site = np.repeat(np.arange(n_sites), rows_per)
    r_lat, r_lon = la[row_site], lo[row_site]

    # obs error
    sig_kind = p["sigma_obs"]
    if sig_kind == "c0":
        sigma = np.full(n_rows, 0.5)
    elif sig_kind == "tuv":
        # sigma = 1 for T, 2 for u and v; the last 3-D variable is T
        sigma = np.where(r_var == three_d[-1], 1.0, 2.0) if three_d else np.full(n_rows, 1.0)
    else:
        sigma = np.exp(rng_o.uniform(np.log(0.3), np.log(3.0), n_rows))
    rinv = 1.0 / sigma ** 2

    # H(x) for truth and members (full fields), then centring
    yb = np.empty((n_rows, kk))
    dep = np.empty(n_rows)
    tflat = truth.reshape(-1)
    mflat = members.reshape(-1, kk)
    noise = rng_o.standard_normal(n_rows) * sigma
    chunk = 1 << 16
    for s in range(0, n_rows, chunk):
        e = min(s + chunk, n_rows)
        idx, w = _h_weights(r_lat[s:e], r_lon[s:e], r_lev[s:e], r_var[s:e], cfg, grid_lat, grid_lon)
        ht = np.einsum('rc,rc->r', tflat[idx], w)
        hx = np.einsum('rck,rc->rk', mflat[idx], w)
        hm = hx.sum(axis=1) / kk
        yb[s:e] = hx - hm[:, None]
        dep[s:e] = (ht + noise[s:e]) - hm

    # background mean / perturbation split (module_da_parallel.F90:195-196)
    h = halo
    axb_i = members.sum(axis=3) / float(np.float32(kk))
    xb_i = members - axb_i[..., None]
    del members
    if h:
        # halo cells hold unrelated values: they are never analysed
        rng_h = _rng(seed, 4)
        xb = rng_h.standard_normal((nlon_ + 2 * h, nlat_ + 2 * h, n, kk))
        axb = rng_h.standard_normal((nlon_ + 2 * h, nlat_ + 2 * h, n))
        xb[h:-h, h:-h] = xb_i
        axb[h:-h, h:-h] = axb_i
    else:
        xb, axb = xb_i, axb_i

    obs = Obs(site_lat=la.astype(np.float64), site_lon=lo.astype(np.float64),
              site_row_ptr=row_ptr.astype(np.int32), yb=yb, dep=dep, rinv=rinv,
              vcoord=r_lev.astype(np.float64), row_var=r_var.astype(np.int32))
    return Instance(cfg=cfg, grid_lat=grid_lat, grid_lon=grid_lon, xb=np.ascontiguousarray(xb),
                    axb=np.ascontiguousarray(axb), truth=truth, obs=obs, name=name, seed=seed)


def full_members(inst: Instance) -> np.ndarray:
    return inst.xb + inst.axb[..., None]


# --------------------------------------------------------------------------
# observation-operator cases (gen_obs_struct inputs)
# --------------------------------------------------------------------------

@dataclass
class ObsOpCase:
    """Inputs of the observation operator: an Instance whose xb/axb carry a
    filled halo of 1 and a hybrid vertical coordinate (cfg.hyai/hybi/p0/ptop,
    ps at state index cfg.ps_index), obs sites with level-major rows, the
    0-based DA variable and level pressure (Pa) of each row, and the observed
    values yo."""
    inst: Instance
    site_lat: np.ndarray
    site_lon: np.ndarray
    site_row_ptr: np.ndarray
    row_var: np.ndarray
    row_ph: np.ndarray
    yo: np.ndarray


def hybrid_coefficients(nz: int, p0: float = 1.0e5, ptop: float = 200.0):
    """Interface coefficients (nz + 1, top to bottom) of a hybrid sigma-pressure
    coordinate: pure pressure near the top, terrain following near the ground."""
    eta = np.linspace(0.0, 1.0, nz + 1)
    hybi = eta ** 2
    hyai = eta - hybi
    return hyai, hybi, p0, ptop


def fill_halo(a: np.ndarray, h: int) -> np.ndarray:
    """Halo of a [lon + 2h, lat + 2h, ...] array: periodic in longitude; past
    the poles, the nearest interior row."""
    nlon = a.shape[0] - 2 * h
    a[:h] = a[nlon:nlon + h]
    a[nlon + h:] = a[h:2 * h]
    a[:, :h] = a[:, h:h + 1]
    a[:, a.shape[1] - h:] = a[:, a.shape[1] - h - 1:a.shape[1] - h]
    return a


def obs_operator_case(name: str = "c2", seed: int = 1, k: Optional[int] = None,
                      nlat: int = 16, nlon: int = 32, nz: int = 6, n_obs: int = 2000,
                      n_stations: int = 20, n_ps: int = 100) -> ObsOpCase:
    """A seeded operator case on config `name`'s variables (ps + 3-D u, v, T):
    single-level obs at random pressures, station columns at every level of a
    sorted pressure profile (all 3-D variables per level), and surface-pressure
    obs.  Positions follow the uniform mix, including the closing cells near
    lat = 90 and lon = 360."""
    inst = generate(name, seed=seed, k=k, nlat=nlat, nlon=nlon, nz=nz, n_obs=8, halo=1)
    cfg = inst.cfg
    hyai, hybi, p0, ptop = hybrid_coefficients(cfg.nz)
    cfg.hyai, cfg.hybi, cfg.p0, cfg.ptop = list(hyai), list(hybi), p0, ptop
    cfg.ps_index = 0
    rng = _rng(seed, 7)
    # ps: ~1e5 Pa mean with a smooth pattern, member spread ~300 Pa
    h = cfg.halo
    lat2 = np.deg2rad(inst.grid_lat)[None, :]
    lon2 = np.deg2rad(inst.grid_lon)[:, None]
    inst.axb[h:-h, h:-h, 0] = 1.0e5 - 3000.0 * np.sin(lat2) ** 2 + 800.0 * np.cos(2 * lon2) * np.cos(lat2)
    inst.xb[h:-h, h:-h, 0, :] *= 300.0
    fill_halo(inst.xb, h)
    fill_halo(inst.axb, h)
    three_d = [v for v in range(cfg.n_var) if cfg.var_lev()[v] > 1]
    two_d = [v for v in range(cfg.n_var) if cfg.var_lev()[v] == 1]
    # sites: single-level obs, station columns, ps obs
    la1, lo1 = _uniform_sphere(rng, n_obs)
    las, los = _uniform_sphere(rng, n_stations)
    lap, lop = _uniform_sphere(rng, n_ps)
    # a few sites in the closing cells (north of the last grid latitude, east of the last longitude)
    m = min(4, n_obs)
    la1[:m] = [89.9, 89.7, -89.9, 10.0][:m]
    lo1[:m] = [359.9, 0.1, 359.95, 359.99][:m]
    row_var, row_ph, rows = [], [], []
    for i in range(n_obs):
        row_var.append(three_d[rng.integers(0, len(three_d))])
        row_ph.append(float(np.exp(rng.uniform(np.log(ptop * 2), np.log(1.05e5)))))
        rows.append(1)
    for i in range(n_stations):
        prof = np.sort(np.exp(rng.uniform(np.log(1000.0), np.log(1.0e5), cfg.nz)))
        for pl in prof:
            for v in three_d:
                row_var.append(v)
                row_ph.append(float(pl))
        rows.append(cfg.nz * len(three_d))
    for i in range(n_ps):
        row_var.append(two_d[0] if two_d else three_d[0])
        row_ph.append(1.0e5)
        rows.append(1)
    site_lat = np.concatenate([la1, las, lap])
    site_lon = np.concatenate([lo1, los, lop])
    rp = np.zeros(len(rows) + 1, np.int32)
    np.cumsum(rows, out=rp[1:])
    yo = rng.standard_normal(int(rp[-1]))
    return ObsOpCase(inst=inst, site_lat=site_lat, site_lon=site_lon, site_row_ptr=rp,
                     row_var=np.asarray(row_var, np.int32), row_ph=np.asarray(row_ph),
                     yo=yo)
