"""LETKF configuration, mirroring the reference's namelist globals.

The fields and defaults follow ``namelist_da`` (x86/DA/utils/da_namelist_mod.F90:27-49)
and the module globals that ``letkf_init`` derives from it
(x86/DA/LETKF/src/letkf_init.F90:27-81):

=====================  ==========================  =========================
field                   reference name              reference default
=====================  ==========================  =========================
local_dist              da_loc_distance (km)        0.0
vert_dist               da_vert_distance            0.0 (<=0: horizontal loop)
inflation_factor        da_inflation_factor         1.0
polar_lat               da_polar_letkf_lat          90.0
gc_filter               da_gc_enable                .false.
gc_filter_vertical      da_gc_enable_vertical       .false.
cube_n                  da_block_interval           4
cube_nz                 da_vertical_interval        32
overlap                 da_block_overlap            2
halo                    da_halo                     3
gamma1, gamma2          gamma1, gamma2              0, 1
precision               computing_precision         53
=====================  ==========================  =========================

``obs_polar`` is ``da_polar_letkf_lat`` when ``polar_discard`` is set and 91
otherwise (letkf_init.F90:50-54).  ``overlap`` = 0 runs the no-overlap loops
(``letkf_compute_nooverlap`` and its vertical variant); 1 .. cube_n - 1 runs
the overlapped-cube ``letkf_compute`` (horizontal loop, FP64 state).
"""
from __future__ import annotations

from dataclasses import dataclass, field, asdict
from typing import List, Optional

import numpy as np

# Precision modes of the engine (the reference selects letkf_core for
# computing_precision >= 52 and letkf_core_mpal below, letkf_compute.F90:105-111).
PREC_FP64 = 0      # obs-space products in FP64 (reference letkf_core)
PREC_FP32 = 1      # FP32 MFMA products, FP64 eigensolve (mpal-style guard)
PREC_BF16X3 = 2    # BF16 hi/lo split MFMA products, FP64 eigensolve

VCOORD_LEVEL = 0   # vertical coordinate is the level index (level units)
VCOORD_LNP = 1     # vertical coordinate is ln p (ps2alnp with hybrid coefficients)


@dataclass
class LetkfConfig:
    k: int                          # ensemble size (svr_ensemble)
    nlat: int                       # reference nx (latitude points of the block)
    nlon: int                       # reference ny (longitude points)
    nz: int = 1                     # num_lev
    n2d: int = 1                    # number of 2-D variables (stored first)
    n3d: int = 0                    # number of 3-D variables (nz levels each)
    halo: int = 0                   # da_halo of the xb/xa arrays
    local_dist: float = 1000.0      # da_loc_distance (km), GC support 2*local_dist
    vert_dist: float = -1.0         # da_vert_distance
    inflation_factor: float = 1.0
    polar_lat: float = 90.0
    polar_discard: bool = False
    gc_filter: bool = True
    gc_filter_vertical: bool = True
    cube_n: int = 1
    cube_nz: int = 1
    overlap: int = 0
    gamma1: float = 0.0
    gamma2: float = 1.0
    precision: int = PREC_FP64
    max_obs: int = 0                # 0: no MAX_OBS cap (the reference caps at 10000)
    vcoord_mode: int = VCOORD_LEVEL
    ps_index: int = 0               # 0-based state index of ps (ps_pos - 1)
    hyai: Optional[List[float]] = None   # nz+1 hybrid coefficients (VCOORD_LNP)
    hybi: Optional[List[float]] = None
    p0: float = 1.0e5
    ptop: float = 0.0

    @property
    def n(self) -> int:
        """State elements per column: 2-D vars first, then nz levels per 3-D var
        (letkf_init.F90:63-73)."""
        return self.n2d + self.n3d * self.nz

    @property
    def obs_polar(self) -> float:
        return self.polar_lat if self.polar_discard else 91.0

    @property
    def n_var(self) -> int:
        return self.n2d + self.n3d

    def var_start(self) -> np.ndarray:
        """1-based start of each variable in n (da_var_start)."""
        out, cur = [], 1
        for _ in range(self.n2d):
            out.append(cur)
            cur += 1
        for _ in range(self.n3d):
            out.append(cur)
            cur += self.nz
        return np.asarray(out, dtype=np.int32)

    def var_lev(self) -> np.ndarray:
        """Levels of each variable (da_var_lev): 1 for 2-D, nz for 3-D."""
        return np.asarray([1] * self.n2d + [self.nz] * self.n3d, dtype=np.int32)

    @property
    def vertical(self) -> bool:
        return self.vert_dist > 0.0

    def n_analyses(self) -> int:
        """One local analysis = one k x k transform solve (one letkf_core call)."""
        cn = max(self.cube_n, 1)
        nh = ((self.nlat + cn - 1) // cn) * ((self.nlon + cn - 1) // cn)
        if self.vertical:
            cz = self.cube_nz if self.cube_nz > 0 else self.nz
            nh *= (self.nz + cz - 1) // cz
        return nh

    def as_dict(self) -> dict:
        return asdict(self)
