"""ctypes binding of the HIP engine (libdida_letkf.so) -- the host-side mirror
of the reference's LETKF call surface:

* :func:`letkf_core` / :func:`letkf_core_batched` -- the external
  ``letkf_core`` (x86/DA/LETKF/src/letkf_core.F90:14) on the GPU;
* :class:`LetkfEngine` -- ``letkf_init`` + ``letkf_compute_nooverlap`` /
  ``letkf_compute_nooverlap_vertical`` (letkf_init.F90:2,
  letkf_compute.F90:12,386) over a grid, an obs set and an ensemble;
* :meth:`LetkfEngine.epilogue_device` -- ``run_da_parallel``'s polar mask,
  gamma blend and spreads (module_da_parallel.F90:219-261).

There is no CPU path: every call goes through the HIP library and raises
if it is missing or fails.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from ._build import LIB_PATH
from .config import LetkfConfig, PREC_FP64

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_longlong)
_vp = C.c_void_p


class DidaConfig(C.Structure):
    _fields_ = [("k", C.c_int), ("nlat", C.c_int), ("nlon", C.c_int), ("nz", C.c_int),
                ("n2d", C.c_int), ("n3d", C.c_int), ("halo", C.c_int), ("cube_n", C.c_int),
                ("cube_nz", C.c_int), ("local_dist", C.c_double), ("vert_dist", C.c_double),
                ("inflation_factor", C.c_double), ("obs_polar", C.c_double),
                ("gc_filter", C.c_int), ("gc_filter_vertical", C.c_int), ("precision", C.c_int),
                ("vcoord_mode", C.c_int), ("ps_index", C.c_int), ("hyai", _dp), ("hybi", _dp),
                ("p0", C.c_double), ("ptop", C.c_double), ("keep_transforms", C.c_int),
                ("max_batch_columns", C.c_int), ("transform_kernel", C.c_int),
                ("products_kernel", C.c_int), ("overlap", C.c_int)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_active", C.c_int)]


class ServeObs(C.Structure):
    _fields_ = [("n_sites", C.c_int), ("site_lat", _vp), ("site_lon", _vp),
                ("site_row_ptr", _vp), ("n_rows", C.c_int), ("yb", _vp), ("dep", _vp),
                ("rinv", _vp), ("vcoord", _vp), ("row_var", _vp), ("row_ph", _vp),
                ("yo", _vp)]


class ServeJob(C.Structure):
    _fields_ = [("cfg", DidaConfig), ("grid_lat", _fp), ("grid_lon", _fp), ("xb", _vp),
                ("axb", _vp), ("n_cycles", C.c_int), ("obs", C.POINTER(ServeObs)),
                ("xa", _vp), ("axa", _vp)]


class ServeStatus(C.Structure):
    _fields_ = [("state", C.c_int), ("rc", C.c_int), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_end", C.c_double), ("n_analyses", C.c_longlong),
                ("bytes", C.c_size_t)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)


class DidaStats(C.Structure):
    _fields_ = [("n_analyses", C.c_longlong), ("total_rows", C.c_longlong),
                ("max_rows", C.c_int), ("n_flagged", C.c_int), ("rows_over_10000", C.c_int),
                ("ms_select", C.c_double), ("ms_products", C.c_double),
                ("ms_transform", C.c_double), ("ms_total", C.c_double),
                ("ms_tridiag", C.c_double), ("ms_ql", C.c_double), ("ms_apply", C.c_double),
                ("ms_update", C.c_double), ("n_rerun", C.c_int),
                ("schedule", C.c_int), ("bucket_analyses", C.c_longlong * 8),
                ("bucket_rows", C.c_longlong * 8), ("bucket_wave_ms", C.c_double * 8),
                ("products_balance", C.c_double), ("variant", C.c_int * 8),
                ("graph_replay", C.c_int), ("reran_rows", C.c_int), ("row_cap", C.c_longlong)]

    def as_dict(self) -> dict:
        out = {}
        for f, _ in self._fields_:
            v = getattr(self, f)
            out[f] = list(v) if isinstance(v, C.Array) else v
        return out


# dida_letkf_set_option
OPT_SCHEDULE, OPT_TIMING, OPT_TUNE, OPT_GRAPH, OPT_ROW_CAP = 1, 2, 3, 4, 5
# length buckets of the local obs count p (upper bounds; the last is open)
BUCKET_EDGES = (0, 256, 512, 1024, 2048, 4096, 8192)


EXPORTS = [
    "dida_letkf_config_default", "dida_letkf_create", "dida_letkf_destroy",
    "dida_letkf_set_stream", "dida_letkf_set_option", "dida_letkf_set_grid",
    "dida_letkf_set_obs",
    "dida_letkf_set_obs_device", "dida_letkf_analyze", "dida_letkf_analyze_device",
    "dida_letkf_analyze_device_f32", "dida_letkf_analyze_epilogue_device",
    "dida_letkf_analyze_epilogue_device_f32", "dida_letkf_split_device",
    "dida_letkf_run_da_device",
    "dida_letkf_get_transforms", "dida_letkf_n_analyses", "dida_letkf_core_batched",
    "dida_letkf_core_batched_noalloc",
    "letkf_core_", "dida_letkf_epilogue_device", "dida_letkf_last_error",
    "dida_letkf_abi_version", "dida_serve_create", "dida_serve_submit", "dida_serve_poll",
    "dida_serve_wait", "dida_serve_cancel", "dida_serve_destroy", "dida_serve_now",
    "dida_letkf_job_bytes", "dida_letkf_obs_operator_device", "dida_letkf_fill_halo_device",
    "dida_obs_records_count", "dida_obs_records_convert",
]

_lib = None


class DidaError(RuntimeError):
    pass


def lib(path: Optional[str] = None):
    """Loads the HIP engine; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise DidaError(f"HIP engine not built: {p} is missing (run __graft_entry__.build())")
    L = C.CDLL(p)
    L.dida_letkf_config_default.argtypes = [C.POINTER(DidaConfig)]
    L.dida_letkf_config_default.restype = None
    L.dida_letkf_create.argtypes = [C.POINTER(DidaConfig), C.POINTER(_vp)]
    L.dida_letkf_destroy.argtypes = [_vp]
    L.dida_letkf_set_stream.argtypes = [_vp, _vp]
    L.dida_letkf_set_option.argtypes = [_vp, C.c_int, C.c_longlong]
    L.dida_letkf_set_grid.argtypes = [_vp, _fp, _fp]
    L.dida_letkf_set_obs.argtypes = [_vp, C.c_int, _dp, _dp, _ip, C.c_int, _dp, _dp, _dp, _dp]
    L.dida_letkf_set_obs_device.argtypes = [_vp, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp,
                                            _vp]
    L.dida_letkf_analyze.argtypes = [_vp, _dp, _dp, _dp, _dp, C.POINTER(DidaStats)]
    L.dida_letkf_analyze_device.argtypes = [_vp, _vp, _vp, _vp, _vp, C.POINTER(DidaStats)]
    L.dida_letkf_analyze_device_f32.argtypes = [_vp, _vp, _vp, _vp, _vp, C.POINTER(DidaStats)]
    L.dida_letkf_analyze_epilogue_device.argtypes = [_vp, _vp, _vp, C.c_double, C.c_double,
                                                     C.c_double, _vp, _vp, _vp,
                                                     C.POINTER(DidaStats)]
    L.dida_letkf_analyze_epilogue_device_f32.argtypes = \
        L.dida_letkf_analyze_epilogue_device.argtypes
    L.dida_letkf_split_device.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.dida_letkf_run_da_device.argtypes = [_vp, _vp, C.c_int, C.c_double, C.c_double,
                                           C.c_double, _vp, _vp, _vp, C.POINTER(DidaStats)]
    L.dida_letkf_get_transforms.argtypes = [_vp, C.c_longlong, _dp, _dp, _ip]
    L.dida_letkf_n_analyses.argtypes = [C.POINTER(DidaConfig)]
    L.dida_letkf_n_analyses.restype = C.c_longlong
    L.dida_letkf_core_batched.argtypes = [C.c_int, C.c_longlong, _lp, _dp, _fp, _dp, _dp, _dp, _dp,
                                          C.c_double, C.c_int, _ip]
    L.dida_letkf_core_batched_noalloc.argtypes = L.dida_letkf_core_batched.argtypes
    L.letkf_core_.argtypes = [_ip, _ip, _dp, _fp, _dp, _dp, _dp, _dp, _dp, _ip]
    L.letkf_core_.restype = None
    L.dida_letkf_epilogue_device.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double,
                                             C.c_double, _vp, _vp, _vp]
    L.dida_letkf_last_error.restype = C.c_char_p
    L.dida_serve_create.argtypes = [C.POINTER(ServeConfig), C.POINTER(_vp)]
    L.dida_serve_submit.argtypes = [_vp, C.POINTER(ServeJob), C.POINTER(C.c_longlong)]
    L.dida_serve_poll.argtypes = [_vp, C.c_longlong, C.POINTER(ServeStatus)]
    L.dida_serve_wait.argtypes = [_vp, C.c_longlong, C.POINTER(ServeStatus)]
    L.dida_serve_cancel.argtypes = [_vp, C.c_longlong]
    L.dida_serve_destroy.argtypes = [_vp]
    L.dida_serve_now.restype = C.c_double
    L.dida_obs_records_count.argtypes = [C.c_int, C.c_int, C.c_int, _fp, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
    L.dida_obs_records_convert.argtypes = [C.c_int, C.c_int, C.c_int, _fp, _fp, _fp, _fp, _fp,
                                           _ip, _dp, _dp, _ip, _ip, _dp, _dp, _dp]
    L.dida_letkf_job_bytes.argtypes = [C.POINTER(DidaConfig), C.c_int, C.c_int]
    L.dida_letkf_job_bytes.restype = C.c_size_t
    _lib = L
    return L


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().dida_letkf_last_error().decode(errors="replace")
        raise DidaError(f"{what} failed ({rc}): {msg}")
    return rc


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def to_c_config(cfg: LetkfConfig, keep_transforms: bool = False,
                max_batch_columns: int = 0, transform_kernel: int = 0,
                products_kernel: int = 0) -> DidaConfig:
    c = DidaConfig()
    lib().dida_letkf_config_default(C.byref(c))
    c.k, c.nlat, c.nlon, c.nz = cfg.k, cfg.nlat, cfg.nlon, cfg.nz
    c.n2d, c.n3d, c.halo = cfg.n2d, cfg.n3d, cfg.halo
    c.cube_n, c.cube_nz = cfg.cube_n, cfg.cube_nz
    c.local_dist, c.vert_dist = cfg.local_dist, cfg.vert_dist
    c.inflation_factor, c.obs_polar = cfg.inflation_factor, cfg.obs_polar
    c.gc_filter, c.gc_filter_vertical = int(cfg.gc_filter), int(cfg.gc_filter_vertical)
    c.precision, c.vcoord_mode, c.ps_index = cfg.precision, cfg.vcoord_mode, cfg.ps_index
    c.p0, c.ptop = cfg.p0, cfg.ptop
    c.keep_transforms = int(keep_transforms)
    c.max_batch_columns = int(max_batch_columns)
    c.transform_kernel = int(transform_kernel)
    c.products_kernel = int(products_kernel)
    c.overlap = int(getattr(cfg, "overlap", 0))
    return c


# ---------------------------------------------------------------------------
# letkf_core
# ---------------------------------------------------------------------------

def letkf_core(hxens: np.ndarray, rdiaginv: np.ndarray, dep: np.ndarray, rloc: np.ndarray,
               inflation_factor: float = 1.0):
    """letkf_core (letkf_core.F90:14) through the by-reference C entry
    ``letkf_core_``.  hxens is (nanals, nobsl).  Returns (trans, overline_w)."""
    k, p = hxens.shape
    hx = np.asfortranarray(hxens, dtype=np.float64)
    rd = np.ascontiguousarray(rdiaginv, dtype=np.float32)
    dp = np.ascontiguousarray(dep, dtype=np.float64)
    rl = np.ascontiguousarray(rloc, dtype=np.float64)
    trans = np.zeros((k, k), order="F")
    w = np.zeros(k)
    kk, pp, pf = C.c_int(k), C.c_int(p), C.c_int(0)
    inf = C.c_double(inflation_factor)
    lib().letkf_core_(C.byref(kk), C.byref(pp), _p(hx, _dp), _p(rd, _fp), _p(dp, _dp),
                      _p(rl, _dp), _p(trans, _dp), _p(w, _dp), C.byref(inf), C.byref(pf))
    return trans, w


def letkf_core_batched(hxens_list, rdiaginv_list, dep_list, rloc_list,
                       inflation_factor: float = 1.0, precision: int = PREC_FP64,
                       noalloc: bool = False):
    """Many independent letkf_core calls in one launch.  Returns
    (trans [n, k, k], overline_w [n, k], status [n]).  noalloc selects the
    entry point that allocates no device memory."""
    n = len(hxens_list)
    k = hxens_list[0].shape[0]
    sizes = np.array([h.shape[1] for h in hxens_list], dtype=np.int64)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(sizes, out=row_ptr[1:])
    tot = int(row_ptr[-1])
    hx = np.zeros((max(tot, 1), k))
    rd = np.zeros(max(tot, 1), np.float32)
    dp = np.zeros(max(tot, 1))
    rl = np.zeros(max(tot, 1))
    for a in range(n):
        s, e = row_ptr[a], row_ptr[a + 1]
        hx[s:e] = np.asarray(hxens_list[a]).T
        rd[s:e] = rdiaginv_list[a]
        dp[s:e] = dep_list[a]
        rl[s:e] = rloc_list[a]
    trans = np.zeros((n, k, k))
    w = np.zeros((n, k))
    st = np.zeros(n, np.int32)
    fn = lib().dida_letkf_core_batched_noalloc if noalloc else lib().dida_letkf_core_batched
    rc = fn(k, n, _p(row_ptr, _lp), _p(hx, _dp), _p(rd, _fp), _p(dp, _dp), _p(rl, _dp),
            _p(trans, _dp), _p(w, _dp), float(inflation_factor), int(precision), _p(st, _ip))
    _check(rc, "dida_letkf_core_batched" + ("_noalloc" if noalloc else ""))
    # stored Fortran (k, k) per analysis: [a, j, i] is element (i, j)
    return np.transpose(trans, (0, 2, 1)).copy(), w, st


# ---------------------------------------------------------------------------
# batched engine
# ---------------------------------------------------------------------------

class LetkfEngine:
    """One handle = one letkf_init'ed block (letkf_init.F90:2)."""

    def __init__(self, cfg: LetkfConfig, keep_transforms: bool = False,
                 max_batch_columns: int = 0, transform_kernel: int = 0,
                 products_kernel: int = 0):
        self.cfg = cfg
        self._keep = []
        c = to_c_config(cfg, keep_transforms, max_batch_columns, transform_kernel,
                        products_kernel)
        if cfg.hyai is not None:
            hy = np.ascontiguousarray(cfg.hyai, np.float64)
            hb = np.ascontiguousarray(cfg.hybi, np.float64)
            self._keep += [hy, hb]
            c.hyai, c.hybi = _p(hy, _dp), _p(hb, _dp)
        self._c = c
        h = _vp()
        _check(lib().dida_letkf_create(C.byref(c), C.byref(h)), "dida_letkf_create")
        self.h = h
        self.n_analyses = int(lib().dida_letkf_n_analyses(C.byref(c)))
        self.stats = DidaStats()

    def close(self):
        if getattr(self, "h", None):
            lib().dida_letkf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_ptr: int):
        _check(lib().dida_letkf_set_stream(self.h, _vp(stream_ptr)), "set_stream")

    def set_option(self, option: int, value: int):
        """dida_letkf_set_option (OPT_SCHEDULE, OPT_TIMING, OPT_TUNE, OPT_GRAPH,
        OPT_ROW_CAP)."""
        _check(lib().dida_letkf_set_option(self.h, int(option), int(value)), "set_option")

    def set_grid(self, grid_lat: np.ndarray, grid_lon: np.ndarray):
        la = np.ascontiguousarray(grid_lat, np.float32)
        lo = np.ascontiguousarray(grid_lon, np.float32)
        _check(lib().dida_letkf_set_grid(self.h, _p(la, _fp), _p(lo, _fp)), "set_grid")

    def set_obs(self, obs):
        """obs: synth.Obs-like (site_lat, site_lon, site_row_ptr, yb, dep, rinv, vcoord)."""
        la = np.ascontiguousarray(obs.site_lat, np.float64)
        lo = np.ascontiguousarray(obs.site_lon, np.float64)
        rp = None if obs.site_row_ptr is None else np.ascontiguousarray(obs.site_row_ptr, np.int32)
        yb = np.ascontiguousarray(obs.yb, np.float64)
        dp = np.ascontiguousarray(obs.dep, np.float64)
        ri = np.ascontiguousarray(obs.rinv, np.float64)
        vc = None if obs.vcoord is None else np.ascontiguousarray(obs.vcoord, np.float64)
        _check(lib().dida_letkf_set_obs(self.h, int(la.size), _p(la, _dp), _p(lo, _dp),
                                        _p(rp, _ip), int(dp.size), _p(yb, _dp), _p(dp, _dp),
                                        _p(ri, _dp), _p(vc, _dp)), "set_obs")

    def set_obs_device(self, n_sites, lat_ptr, lon_ptr, rowptr_ptr, n_rows, yb_ptr, dep_ptr,
                       rinv_ptr, vco_ptr):
        _check(lib().dida_letkf_set_obs_device(self.h, int(n_sites), _vp(lat_ptr), _vp(lon_ptr),
                                               _vp(rowptr_ptr) if rowptr_ptr else None,
                                               int(n_rows), _vp(yb_ptr), _vp(dep_ptr),
                                               _vp(rinv_ptr), _vp(vco_ptr) if vco_ptr else None),
               "set_obs_device")

    def letkf_compute(self, xb: np.ndarray, axb: np.ndarray, xa: Optional[np.ndarray] = None,
                      axa: Optional[np.ndarray] = None):
        """letkf_compute_nooverlap[_vertical] on host arrays; returns (xa, axa, n_flagged)."""
        xb = np.ascontiguousarray(xb, np.float64)
        axb = np.ascontiguousarray(axb, np.float64)
        xa = np.zeros_like(xb) if xa is None else xa
        axa = np.zeros_like(axb) if axa is None else axa
        rc = lib().dida_letkf_analyze(self.h, _p(xb, _dp), _p(axb, _dp), _p(xa, _dp),
                                      _p(axa, _dp), C.byref(self.stats))
        _check(rc, "dida_letkf_analyze")
        return xa, axa, rc

    def analyze_device(self, xb_ptr: int, axb_ptr: int, xa_ptr: int, axa_ptr: int) -> int:
        rc = lib().dida_letkf_analyze_device(self.h, _vp(xb_ptr), _vp(axb_ptr), _vp(xa_ptr),
                                             _vp(axa_ptr), C.byref(self.stats))
        return _check(rc, "dida_letkf_analyze_device")

    def analyze_device_f32(self, xb_ptr: int, axb_ptr: int, xa_ptr: int, axa_ptr: int) -> int:
        """The analysis on an FP32 perturbation state (FP64 means)."""
        rc = lib().dida_letkf_analyze_device_f32(self.h, _vp(xb_ptr), _vp(axb_ptr), _vp(xa_ptr),
                                                 _vp(axa_ptr), C.byref(self.stats))
        return _check(rc, "dida_letkf_analyze_device_f32")

    def analyze_epilogue_device(self, xb_ptr: int, axb_ptr: int, polar_lat: float,
                                gamma1: float, gamma2: float, xa_out_ptr: int, sb_ptr: int,
                                sa_ptr: int, f32: bool = False) -> int:
        """The analysis with run_da_parallel's epilogue fused into the update
        (module_da_parallel.F90:219-261); f32: the perturbations are FP32."""
        fn = (lib().dida_letkf_analyze_epilogue_device_f32 if f32
              else lib().dida_letkf_analyze_epilogue_device)
        rc = fn(self.h, _vp(xb_ptr), _vp(axb_ptr), float(polar_lat), float(gamma1),
                float(gamma2), _vp(xa_out_ptr), _vp(sb_ptr), _vp(sa_ptr), C.byref(self.stats))
        return _check(rc, "dida_letkf_analyze_epilogue_device" + ("_f32" if f32 else ""))

    def split_device(self, x_in_ptr: int, xb_ptr: int = 0, xb32_ptr: int = 0,
                     axb_ptr: int = 0) -> int:
        """dnn2letkf on the device (module_da_parallel.F90:186-211): mean and
        perturbations (FP64 into xb, or FP32 into xb32) of the full state."""
        return _check(lib().dida_letkf_split_device(self.h, _vp(x_in_ptr), _vp(xb_ptr or None),
                                                    _vp(xb32_ptr or None), _vp(axb_ptr)),
                      "dida_letkf_split_device")

    def run_da_device(self, x_in_ptr: int, polar_lat: float, gamma1: float, gamma2: float,
                      xa_out_ptr: int, sb_ptr: int, sa_ptr: int, state_f32: bool = False) -> int:
        """run_da_parallel's analysis step (module_da_parallel.F90:186-261) on
        the device: split, analysis, fused letkf2dnn epilogue."""
        rc = lib().dida_letkf_run_da_device(self.h, _vp(x_in_ptr), int(bool(state_f32)),
                                            float(polar_lat), float(gamma1), float(gamma2),
                                            _vp(xa_out_ptr), _vp(sb_ptr), _vp(sa_ptr),
                                            C.byref(self.stats))
        return _check(rc, "dida_letkf_run_da_device")

    def get_transforms(self):
        k = self.cfg.k
        n = self.n_analyses
        trans = np.zeros((n, k, k))
        w = np.zeros((n, k))
        nob = np.zeros(n, np.int32)
        _check(lib().dida_letkf_get_transforms(self.h, n, _p(trans, _dp), _p(w, _dp),
                                               _p(nob, _ip)), "get_transforms")
        return np.transpose(trans, (0, 2, 1)).copy(), w, nob

    def local_obs_counts(self):
        """Local obs rows of each analysis of the last analyze call (needs
        keep_transforms)."""
        nob = np.zeros(self.n_analyses, np.int32)
        _check(lib().dida_letkf_get_transforms(self.h, self.n_analyses, None, None,
                                               _p(nob, _ip)), "get_transforms")
        return nob

    def obs_operator_device(self, xb_ptr, axb_ptr, n_sites, lat_ptr, lon_ptr, row_ptr_ptr,
                            n_rows, row_var_ptr, row_ph_ptr, yo_ptr, yb_ptr, dep_ptr):
        """H(x) and Yb/d centring (gen_obs_struct) on device arrays."""
        _check(lib().dida_letkf_obs_operator_device(
            self.h, _vp(xb_ptr), _vp(axb_ptr), int(n_sites), _vp(lat_ptr), _vp(lon_ptr),
            _vp(row_ptr_ptr), int(n_rows), _vp(row_var_ptr), _vp(row_ph_ptr), _vp(yo_ptr),
            _vp(yb_ptr), _vp(dep_ptr)), "obs_operator")

    def fill_halo_device(self, x_ptr, ax_ptr):
        _check(lib().dida_letkf_fill_halo_device(self.h, _vp(x_ptr), _vp(ax_ptr)), "fill_halo")

    def epilogue_device(self, xb_ptr, axb_ptr, xa_ptr, axa_ptr, polar_lat, gamma1, gamma2,
                        xa_out_ptr, sb_ptr, sa_ptr):
        _check(lib().dida_letkf_epilogue_device(self.h, _vp(xb_ptr), _vp(axb_ptr), _vp(xa_ptr),
                                                _vp(axa_ptr), float(polar_lat), float(gamma1),
                                                float(gamma2), _vp(xa_out_ptr), _vp(sb_ptr),
                                                _vp(sa_ptr)), "epilogue")


def obs_operator(case):
    """Convenience: the device observation operator on a synth.ObsOpCase
    (inputs copied to the GPU).  Returns (yb [n_rows, k], dep [n_rows])."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).to(dev)
    inst = case.inst
    cfg = inst.cfg
    nr = int(case.site_row_ptr[-1])
    xb, axb = t(inst.xb), t(inst.axb)
    la, lo, rp = t(case.site_lat, np.float64), t(case.site_lon, np.float64), t(case.site_row_ptr, np.int32)
    rv, rph, yo = t(case.row_var, np.int32), t(case.row_ph, np.float64), t(case.yo, np.float64)
    yb = torch.empty((max(nr, 1), cfg.k), dtype=torch.float64, device=dev)
    dep = torch.empty(max(nr, 1), dtype=torch.float64, device=dev)
    with LetkfEngine(cfg) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.obs_operator_device(xb.data_ptr(), axb.data_ptr(), case.site_lat.size, la.data_ptr(),
                                lo.data_ptr(), rp.data_ptr(), nr, rv.data_ptr(), rph.data_ptr(),
                                yo.data_ptr(), yb.data_ptr(), dep.data_ptr())
        torch.cuda.synchronize()
    return yb[:nr].cpu().numpy(), dep[:nr].cpu().numpy()


def letkf_compute_nooverlap(inst, precision: Optional[int] = None, keep_transforms=False,
                            transform_kernel: int = 0, products_kernel: int = 0):
    """Convenience: run one analysis cycle of a synth.Instance on the GPU."""
    cfg = inst.cfg
    if precision is not None:
        import copy
        cfg = copy.copy(cfg)
        cfg.precision = precision
    with LetkfEngine(cfg, keep_transforms=keep_transforms, transform_kernel=transform_kernel,
                     products_kernel=products_kernel) as eng:
        eng.set_grid(inst.grid_lat, inst.grid_lon)
        eng.set_obs(inst.obs)
        xa, axa, nflag = eng.letkf_compute(inst.xb, inst.axb)
        out = dict(xa=xa, axa=axa, n_flagged=nflag, stats=eng.stats.as_dict())
        if keep_transforms:
            t, w, nob = eng.get_transforms()
            out.update(trans=t, overline_w=w, nobsl=nob)
        return out


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------

class Server:
    """Continuous-batching server (dida_serve_*): FIFO admission by HBM budget,
    one stream per running job, device-resident ensembles across cycles."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 4):
        c = ServeConfig(hbm_budget_bytes=int(hbm_budget_bytes), max_active=int(max_active))
        h = _vp()
        _check(lib().dida_serve_create(C.byref(c), C.byref(h)), "dida_serve_create")
        self.h = h
        self._keep = {}

    def submit(self, cfg: LetkfConfig, grid_lat, grid_lon, xb_ptr, axb_ptr, obs_cycles,
               xa_ptr, axa_ptr, max_batch_columns: int = 0) -> int:
        """obs_cycles: list of dicts with device pointers n_sites, site_lat, site_lon,
        site_row_ptr, n_rows, rinv, vcoord and either yb, dep or (operator mode)
        row_var, row_ph, yo.  max_batch_columns: 0 lets the server bound each
        batch's scratch to 4 GiB."""
        job = ServeJob()
        job.cfg = to_c_config(cfg, max_batch_columns=max_batch_columns)
        if cfg.hyai is not None:          # copied by the server at submit
            hy = np.ascontiguousarray(cfg.hyai, np.float64)
            hb = np.ascontiguousarray(cfg.hybi, np.float64)
            job.cfg.hyai, job.cfg.hybi = _p(hy, _dp), _p(hb, _dp)
        la = np.ascontiguousarray(grid_lat, np.float32)
        lo = np.ascontiguousarray(grid_lon, np.float32)
        arr = (ServeObs * len(obs_cycles))()
        for i, o in enumerate(obs_cycles):
            arr[i] = ServeObs(n_sites=int(o["n_sites"]), site_lat=o["site_lat"],
                              site_lon=o["site_lon"], site_row_ptr=o.get("site_row_ptr"),
                              n_rows=int(o["n_rows"]), yb=o.get("yb"), dep=o.get("dep"),
                              rinv=o["rinv"], vcoord=o.get("vcoord"), row_var=o.get("row_var"),
                              row_ph=o.get("row_ph"), yo=o.get("yo"))
        job.grid_lat, job.grid_lon = _p(la, _fp), _p(lo, _fp)
        job.xb, job.axb, job.xa, job.axa = xb_ptr, axb_ptr, xa_ptr, axa_ptr
        job.n_cycles = len(obs_cycles)
        job.obs = C.cast(arr, C.POINTER(ServeObs))
        jid = C.c_longlong()
        _check(lib().dida_serve_submit(self.h, C.byref(job), C.byref(jid)), "dida_serve_submit")
        return int(jid.value)

    def poll(self, jid: int) -> dict:
        st = ServeStatus()
        _check(lib().dida_serve_poll(self.h, jid, C.byref(st)), "dida_serve_poll")
        return st.as_dict()

    def wait(self, jid: int) -> dict:
        st = ServeStatus()
        _check(lib().dida_serve_wait(self.h, jid, C.byref(st)), "dida_serve_wait")
        return st.as_dict()

    def cancel(self, jid: int):
        _check(lib().dida_serve_cancel(self.h, jid), "dida_serve_cancel")

    @staticmethod
    def now() -> float:
        return lib().dida_serve_now()

    def close(self):
        if getattr(self, "h", None):
            lib().dida_serve_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_bytes(cfg: LetkfConfig, n_sites: int, n_rows: int) -> int:
    c = to_c_config(cfg)
    return int(lib().dida_letkf_job_bytes(C.byref(c), int(n_sites), int(n_rows)))
