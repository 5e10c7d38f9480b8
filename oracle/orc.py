"""ctypes binding of the CPU oracle (liborc_letkf.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc_letkf.so")

ORC_EIG_JACOBI = 1
ORC_GUARD = 2

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class OrcObs(C.Structure):
    _fields_ = [("n_sites", C.c_int), ("n_rows", C.c_int),
                ("site_lat", _dp), ("site_lon", _dp), ("site_row_ptr", _ip),
                ("yb", _dp), ("dep", _dp), ("rinv", _dp), ("vcoord", _dp)]


class OrcProblem(C.Structure):
    _fields_ = [("k", C.c_int), ("n", C.c_int), ("nlat", C.c_int), ("nlon", C.c_int),
                ("halo", C.c_int), ("nz", C.c_int), ("cube_n", C.c_int), ("cube_nz", C.c_int),
                ("grid_lat", _fp), ("grid_lon", _fp),
                ("local_dist", C.c_double), ("vert_dist", C.c_double),
                ("inflation_factor", C.c_double), ("obs_polar", C.c_double),
                ("gc_filter", C.c_int), ("gc_filter_vertical", C.c_int), ("max_obs", C.c_int),
                ("core_flags", C.c_int), ("vcoord_mode", C.c_int), ("ps_index", C.c_int),
                ("n_var", C.c_int), ("var_start", _ip), ("var_lev", _ip),
                ("hyai", _dp), ("hybi", _dp), ("p0", C.c_double), ("ptop", C.c_double),
                ("obs", C.POINTER(OrcObs)), ("overlap", C.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_schr2.restype = C.c_double
        L.orc_schr2.argtypes = [C.c_double]
        L.orc_schr.restype = C.c_double
        L.orc_schr.argtypes = [C.c_double]
        L.orc_pi_f32.restype = C.c_double
        L.orc_earth_radius_km.restype = C.c_double
        L.orc_latlon_dist.restype = C.c_double
        L.orc_latlon_dist.argtypes = [C.c_double] * 4
        L.orc_syev_ql.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_syev_jacobi.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_letkf_core.argtypes = [C.c_int, C.c_int, _dp, _fp, _dp, _dp, _dp, _dp, C.c_double,
                                     C.c_int]
        L.orc_letkf_transform.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
        L.orc_xb_2_xa.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_n_analyses.argtypes = [C.POINTER(OrcProblem)]
        L.orc_letkf_compute.argtypes = [C.POINTER(OrcProblem), _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                        _ip, _ip, C.c_int, C.c_int]
        L.orc_epilogue.argtypes = [C.POINTER(OrcProblem), _dp, _dp, _dp, _dp, C.c_double,
                                   C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_baseline_run.argtypes = [C.POINTER(OrcProblem), _dp, _dp, _dp, _dp, C.c_int,
                                       C.c_int, C.POINTER(C.c_longlong),
                                       C.POINTER(C.c_longlong), C.POINTER(C.c_double), _ip]
        L.orc_baseline_provider.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def schr2(ri: float) -> float:
    return lib().orc_schr2(float(ri))


def schr(ri: float) -> float:
    return lib().orc_schr(float(ri))


def latlon_dist(x1, y1, x2, y2) -> float:
    return lib().orc_latlon_dist(float(x1), float(y1), float(x2), float(y2))


def syev(a: np.ndarray, method: str = "ql"):
    """Symmetric eigendecomposition (ascending) of a (n x n, Fortran order)."""
    n = a.shape[0]
    a = np.asfortranarray(a, dtype=np.float64)
    w = np.empty(n)
    v = np.empty((n, n), order="F")
    fn = lib().orc_syev_ql if method == "ql" else lib().orc_syev_jacobi
    info = fn(n, _p(a, _dp), _p(w, _dp), _p(v, _dp))
    return w, v, info


def letkf_core(hxens: np.ndarray, rdiaginv: np.ndarray, dep: np.ndarray, rloc: np.ndarray,
               inflation_factor: float = 1.0, jacobi: bool = False, guard: bool = False):
    """letkf_core.F90:14 -- hxens is (nanals, nobsl). Returns (trans, overline_w, info)."""
    k, p = hxens.shape
    hx = np.asfortranarray(hxens, dtype=np.float64)
    rd = np.ascontiguousarray(rdiaginv, dtype=np.float32)
    dp = np.ascontiguousarray(dep, dtype=np.float64)
    rl = np.ascontiguousarray(rloc, dtype=np.float64)
    trans = np.empty((k, k), order="F")
    w = np.empty(k)
    flags = (ORC_EIG_JACOBI if jacobi else 0) | (ORC_GUARD if guard else 0)
    info = lib().orc_letkf_core(k, p, _p(hx, _dp), _p(rd, _fp), _p(dp, _dp), _p(rl, _dp),
                                _p(trans, _dp), _p(w, _dp), float(inflation_factor), flags)
    return trans, w, info


def letkf_transform(a: np.ndarray, b: np.ndarray, inflation_factor: float = 1.0,
                    jacobi: bool = False, guard: bool = False):
    """Transform half of letkf_core from A = Yb^T R^-1 Yb and b = Yb^T R^-1 d."""
    k = a.shape[0]
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    trans = np.empty((k, k), order="F")
    w = np.empty(k)
    flags = (ORC_EIG_JACOBI if jacobi else 0) | (ORC_GUARD if guard else 0)
    info = lib().orc_letkf_transform(k, _p(a, _dp), _p(b, _dp), float(inflation_factor), flags,
                                     _p(trans, _dp), _p(w, _dp))
    return trans, w, info


def xb_2_xa(xb: np.ndarray, axb: np.ndarray, trans: np.ndarray):
    """xb_2_xa for a block: xb (m, k) C-order (k fastest), axb (m,)."""
    m, k = xb.shape
    xb = np.ascontiguousarray(xb, dtype=np.float64)
    axb = np.ascontiguousarray(axb, dtype=np.float64)
    tr = np.asfortranarray(trans, dtype=np.float64)
    xa = np.empty_like(xb)
    axa = np.empty(m)
    lib().orc_xb_2_xa(k, m, _p(xb, _dp), _p(axb, _dp), _p(xa, _dp), _p(axa, _dp), _p(tr, _dp))
    return xa, axa


class _Keep:
    """Holds numpy buffers alive while a ctypes structure points at them."""

    def __init__(self):
        self.refs = []

    def __call__(self, a):
        self.refs.append(a)
        return a


def _problem(inst, core_flags: int = 0):
    cfg = inst.cfg
    o = inst.obs
    keep = _Keep()
    obs = OrcObs()
    obs.n_sites = o.n_sites
    obs.n_rows = o.n_rows
    obs.site_lat = _p(keep(np.ascontiguousarray(o.site_lat, np.float64)), _dp)
    obs.site_lon = _p(keep(np.ascontiguousarray(o.site_lon, np.float64)), _dp)
    obs.site_row_ptr = _p(keep(np.ascontiguousarray(o.site_row_ptr, np.int32)), _ip)
    obs.yb = _p(keep(np.ascontiguousarray(o.yb, np.float64)), _dp)
    obs.dep = _p(keep(np.ascontiguousarray(o.dep, np.float64)), _dp)
    obs.rinv = _p(keep(np.ascontiguousarray(o.rinv, np.float64)), _dp)
    obs.vcoord = _p(keep(np.ascontiguousarray(o.vcoord, np.float64)), _dp)
    keep(obs)
    pb = OrcProblem()
    pb.k, pb.n, pb.nlat, pb.nlon = cfg.k, cfg.n, cfg.nlat, cfg.nlon
    pb.halo, pb.nz, pb.cube_n, pb.cube_nz = cfg.halo, cfg.nz, cfg.cube_n, cfg.cube_nz
    pb.grid_lat = _p(keep(np.ascontiguousarray(inst.grid_lat, np.float32)), _fp)
    pb.grid_lon = _p(keep(np.ascontiguousarray(inst.grid_lon, np.float32)), _fp)
    pb.local_dist = cfg.local_dist
    pb.vert_dist = cfg.vert_dist
    pb.inflation_factor = cfg.inflation_factor
    pb.obs_polar = cfg.obs_polar
    pb.gc_filter = int(cfg.gc_filter)
    pb.gc_filter_vertical = int(cfg.gc_filter_vertical)
    pb.max_obs = cfg.max_obs
    pb.core_flags = core_flags
    pb.vcoord_mode = cfg.vcoord_mode
    pb.ps_index = cfg.ps_index
    pb.n_var = cfg.n_var
    pb.var_start = _p(keep(cfg.var_start()), _ip)
    pb.var_lev = _p(keep(cfg.var_lev()), _ip)
    if cfg.hyai is not None:
        pb.hyai = _p(keep(np.asarray(cfg.hyai, np.float64)), _dp)
        pb.hybi = _p(keep(np.asarray(cfg.hybi, np.float64)), _dp)
    pb.p0, pb.ptop = cfg.p0, cfg.ptop
    pb.obs = C.pointer(obs)
    pb.overlap = int(getattr(cfg, "overlap", 0))
    return pb, keep


def letkf_compute(inst, nthreads: Optional[int] = None, columns: Optional[np.ndarray] = None,
                  want_trans: bool = False, jacobi: bool = False, guard: bool = False):
    """letkf_compute_nooverlap / _vertical on an Instance, or letkf_compute
    (overlapped cubes, every column; no column subset) when cfg.overlap > 0.

    Returns dict with xa, axa (same shapes as inst.xb / inst.axb; only the
    analysed columns are written, the rest are zero), and optionally
    trans (n_analyses, k, k) column-major per analysis, overline_w, nobsl.
    columns: optional array of horizontal analysis indices (lon-major).
    """
    cfg = inst.cfg
    pb, keep = _problem(inst, (ORC_EIG_JACOBI if jacobi else 0) | (ORC_GUARD if guard else 0))
    L = lib()
    na = L.orc_n_analyses(C.byref(pb))
    xb = np.ascontiguousarray(inst.xb, np.float64)
    axb = np.ascontiguousarray(inst.axb, np.float64)
    xa = np.zeros_like(xb)
    axa = np.zeros_like(axb)
    trans = np.zeros((na, cfg.k, cfg.k)) if want_trans else None
    w = np.zeros((na, cfg.k)) if want_trans else None
    nobsl = np.full(na, -1, np.int32)
    status = np.zeros(na, np.int32)
    cols = None if columns is None else np.ascontiguousarray(columns, np.int32)
    if cols is not None and cols.size:
        nh = (-(-cfg.nlat // cfg.cube_n)) * (-(-cfg.nlon // cfg.cube_n))
        if int(cols.min()) < 0 or int(cols.max()) >= nh:
            raise ValueError(f"columns must be horizontal indices in [0, {nh})")
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = L.orc_letkf_compute(C.byref(pb), _p(xb, _dp), _p(axb, _dp), _p(xa, _dp), _p(axa, _dp),
                             _p(trans, _dp), _p(w, _dp), _p(nobsl, _ip), _p(status, _ip),
                             _p(cols, _ip), 0 if cols is None else int(cols.size), int(nthreads))
    if rc == -3:
        raise ValueError("overlapped cubes: horizontal loop only, every column, "
                         "overlap < cube_n")
    if rc:
        raise RuntimeError(f"oracle letkf_compute failed rc={rc}")
    out = dict(xa=xa, axa=axa, nobsl=nobsl, status=status)
    if want_trans:
        # stored per analysis as Fortran (k, k): element (i, j) at [a, j, i]
        out["trans"] = np.transpose(trans, (0, 2, 1)).copy()
        out["overline_w"] = w
    return out


def epilogue(inst, xa, axa, polar_lat=None, gamma1=None, gamma2=None):
    cfg = inst.cfg
    pb, keep = _problem(inst)
    pl = cfg.polar_lat if polar_lat is None else polar_lat
    g1 = cfg.gamma1 if gamma1 is None else gamma1
    g2 = cfg.gamma2 if gamma2 is None else gamma2
    xa_out = np.zeros_like(inst.xb)
    sb = np.zeros((cfg.nlon, cfg.nlat, cfg.n))
    sa = np.zeros((cfg.nlon, cfg.nlat, cfg.n))
    xb = np.ascontiguousarray(inst.xb)
    axb = np.ascontiguousarray(inst.axb)
    xa = np.ascontiguousarray(xa)
    axa = np.ascontiguousarray(axa)
    lib().orc_epilogue(C.byref(pb), _p(xb, _dp), _p(axb, _dp), _p(xa, _dp), _p(axa, _dp),
                       float(pl), float(g1), float(g2), _p(xa_out, _dp), _p(sb, _dp), _p(sa, _dp))
    return xa_out, sb, sa


def obs_operator(case):
    """gen_obs_struct's H(x) and centring on an ObsOpCase (synth.obs_operator_case).
    Returns (yb [n_rows, k], dep [n_rows])."""
    inst = case.inst
    pb, keep = _problem(inst)
    xb = np.ascontiguousarray(inst.xb, np.float64)
    axb = np.ascontiguousarray(inst.axb, np.float64)
    o = case
    nr = int(o.site_row_ptr[-1])
    yb = np.zeros((nr, inst.cfg.k))
    dep = np.zeros(nr)
    la = keep(np.ascontiguousarray(o.site_lat, np.float64))
    lo = keep(np.ascontiguousarray(o.site_lon, np.float64))
    rp = keep(np.ascontiguousarray(o.site_row_ptr, np.int32))
    rv = keep(np.ascontiguousarray(o.row_var, np.int32))
    rph = keep(np.ascontiguousarray(o.row_ph, np.float64))
    yo = keep(np.ascontiguousarray(o.yo, np.float64))
    rc = lib().orc_obs_operator(C.byref(pb), _p(xb, _dp), _p(axb, _dp), int(la.size), _p(la, _dp),
                                _p(lo, _dp), _p(rp, _ip), _p(rv, _ip), _p(rph, _dp), _p(yo, _dp),
                                _p(yb, _dp), _p(dep, _dp))
    if rc != 0:
        raise ValueError("orc_obs_operator: bad arguments")
    return yb, dep


def baseline(inst, nthreads: int, sample: int = 1):
    """The CPU baseline of bench.py (letkf_baseline.c): the horizontal loop on
    nthreads ranks of the reference's block split with their comm_info_init
    candidate obs, letkf_core/xb_2_xa on MKL when it loads, every sample-th
    column of each block.  Returns dict(xa, axa, n_analyses, n_candidates,
    wall_s, split, provider)."""
    pb, keep = _problem(inst)
    xb = np.ascontiguousarray(inst.xb, np.float64)
    axb = np.ascontiguousarray(inst.axb, np.float64)
    xa = np.zeros_like(xb)
    axa = np.zeros_like(axb)
    na, nc, wall = C.c_longlong(), C.c_longlong(), C.c_double()
    split = np.zeros(2, np.int32)
    rc = lib().orc_baseline_run(C.byref(pb), _p(xb, _dp), _p(axb, _dp), _p(xa, _dp),
                                _p(axa, _dp), int(nthreads), int(sample), C.byref(na),
                                C.byref(nc), C.byref(wall), _p(split, _ip))
    if rc != 0:
        raise ValueError("orc_baseline_run: unsupported problem (vertical loop or cube_n > 1)")
    return dict(xa=xa, axa=axa, n_analyses=int(na.value), n_candidates=int(nc.value),
                wall_s=float(wall.value), split=(int(split[0]), int(split[1])),
                provider=lib().orc_baseline_provider().decode())
