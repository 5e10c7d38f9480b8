"""Observation ingestion: the records read_obs receives from its store
(read_obs_mod.F90:338-449, RedisGetobserveF's buffers) into the sites and
rows of the observation operator and set_obs, through the C ABI
(``dida_obs_records_count`` / ``dida_obs_records_convert``).

File dump of one obs type's records (little-endian), written by the caller
in place of the Redis read::

    8 bytes   b"DIDAOBS1"
    int32     nvar, multi_level (1: ADPUPA / GPSRO profiles), n_records
    float32   lon[n_records], lat[n_records], lev[n_records]
    float32   val[n_records][2 * nvar + 1]   (value_1, oe_1, ..., value_nvar,
                                              oe_nvar, ph) per record

The values, R^-1 = 1/oe**2 and the perturbed value + noise * oe are real(4)
as in the reference; the standard normal draws that read_obs takes from its
randn come from the caller (none by default).
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

MAGIC = b"DIDAOBS1"


@dataclass
class ObsRecords:
    nvar: int
    multi_level: bool
    lon: np.ndarray      # float32 (n_records,)
    lat: np.ndarray
    lev: np.ndarray
    val: np.ndarray      # float32 (n_records, 2 nvar + 1)

    @property
    def n_records(self) -> int:
        return int(self.lon.size)


def write_records(path: str, rec: ObsRecords) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<iii", rec.nvar, int(bool(rec.multi_level)), rec.n_records))
        for a in (rec.lon, rec.lat, rec.lev, rec.val):
            f.write(np.ascontiguousarray(a, "<f4").tobytes())


def read_records(path: str) -> ObsRecords:
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != MAGIC:
        raise ValueError(f"{path}: not an obs record dump")
    nvar, ml, n = struct.unpack_from("<iii", data, 8)
    if nvar < 1 or n < 0:
        raise ValueError(f"{path}: bad header nvar={nvar} n_records={n}")
    w = 2 * nvar + 1
    need = 20 + 4 * n * (3 + w)
    if len(data) != need:
        raise ValueError(f"{path}: {len(data)} bytes, the header says {need}")
    arr = np.frombuffer(data, "<f4", offset=20)
    lon, lat, lev = arr[:n], arr[n:2 * n], arr[2 * n:3 * n]
    val = arr[3 * n:].reshape(n, w)
    return ObsRecords(nvar, bool(ml), lon.astype(np.float32), lat.astype(np.float32),
                      lev.astype(np.float32), val.astype(np.float32))


def to_rows(rec: ObsRecords, var_index: Sequence[int], noise: Optional[np.ndarray] = None):
    """Sites and rows (dida_obs_records_convert).  var_index[n]: 0-based DA
    variable of obs variable n.  Returns a dict with site_lat, site_lon,
    site_row_ptr, row_var, row_ph, yo, rinv."""
    from .engine import lib, _check, _p, _dp, _fp, _ip
    L = lib()
    n, nv = rec.n_records, rec.nvar
    if len(var_index) != nv:
        raise ValueError("var_index needs one DA variable per obs variable")
    lon = np.ascontiguousarray(rec.lon, np.float32)
    lat = np.ascontiguousarray(rec.lat, np.float32)
    lev = np.ascontiguousarray(rec.lev, np.float32)
    val = np.ascontiguousarray(rec.val, np.float32).reshape(n, 2 * nv + 1)
    vi = np.ascontiguousarray(var_index, np.int32)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float32).reshape(n, nv)
    ns, nr = C.c_int(), C.c_int()
    _check(L.dida_obs_records_count(n, nv, int(rec.multi_level), _p(lev, _fp), C.byref(ns),
                                    C.byref(nr)), "dida_obs_records_count")
    out = dict(site_lat=np.zeros(ns.value), site_lon=np.zeros(ns.value),
               site_row_ptr=np.zeros(ns.value + 1, np.int32), row_var=np.zeros(nr.value, np.int32),
               row_ph=np.zeros(nr.value), yo=np.zeros(nr.value), rinv=np.zeros(nr.value))
    _check(L.dida_obs_records_convert(
        n, nv, int(rec.multi_level), _p(lon, _fp), _p(lat, _fp), _p(lev, _fp), _p(val, _fp),
        _p(nz, _fp), _p(vi, _ip), _p(out["site_lat"], _dp), _p(out["site_lon"], _dp),
        _p(out["site_row_ptr"], _ip), _p(out["row_var"], _ip), _p(out["row_ph"], _dp),
        _p(out["yo"], _dp), _p(out["rinv"], _dp)), "dida_obs_records_convert")
    return out
