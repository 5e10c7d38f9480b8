"""Writes the obs record fixtures (the layout read_obs receives,
read_obs_mod.F90:338-449; format in dida_letkf/obs_records.py):

* obs_records_profiles.bin: multi-level (ADPUPA-like) records of 12 station
  profiles, variables u, v, T, levels 1..5 at decreasing pressure; station 4
  lacks level 3 and station 9 has only level 1;
* obs_records_surface.bin: single-level records of 30 surface-pressure obs.

Synthetic, seeded: the reference reads its records from Redis, which is not
part of this repository.  Run from the repository root:
python tests/golden/make_obs_records.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "dida-thhpca_amd"))

from dida_letkf.obs_records import ObsRecords, write_records  # noqa: E402


def profiles(rng):
    lon, lat, lev, val = [], [], [], []
    for s in range(12):
        lo, la = rng.uniform(0.0, 360.0), rng.uniform(-70.0, 70.0)
        levels = [1] if s == 9 else [l for l in range(1, 6) if not (s == 4 and l == 3)]
        for l in levels:
            ph = 95000.0 - 14000.0 * (l - 1) + rng.uniform(-500.0, 500.0)
            row = []
            for scale in (8.0, 8.0, 15.0):              # u, v, T
                row += [rng.normal(0.0, scale), rng.uniform(0.4, 3.0)]
            row.append(ph)
            lon.append(lo), lat.append(la), lev.append(float(l)), val.append(row)
    f = lambda a: np.asarray(a, np.float32)
    return ObsRecords(3, True, f(lon), f(lat), f(lev), f(val))


def surface(rng):
    n = 30
    ps = rng.uniform(97000.0, 103000.0, n)
    val = np.stack([ps, rng.uniform(50.0, 150.0, n), ps], axis=1)
    return ObsRecords(1, False, rng.uniform(0.0, 360.0, n).astype(np.float32),
                      rng.uniform(-80.0, 80.0, n).astype(np.float32),
                      np.ones(n, np.float32), val.astype(np.float32))


if __name__ == "__main__":
    rng = np.random.default_rng(20261203)
    write_records(os.path.join(HERE, "obs_records_profiles.bin"), profiles(rng))
    write_records(os.path.join(HERE, "obs_records_surface.bin"), surface(rng))
    print("wrote obs_records_profiles.bin, obs_records_surface.bin")
