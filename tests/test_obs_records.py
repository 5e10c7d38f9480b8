"""Obs ingestion (dida_obs_records_count / _convert): the records read_obs
receives (read_obs_mod.F90:338-449) into sites and rows, on the committed
fixtures (tests/golden/obs_records_*.bin, written by make_obs_records.py).

Checked against a restatement of read_obs's assignment loops in numpy
(:406-442: a profile site starts at lev == 1 and takes the records after it;
every variable of a record is a row; value + randn * oe and 1 / oe**2 in
real(4)), against closed forms (rinv * oe**2 = 1 within a real(4) rounding;
rows = nvar per record; yo = value without noise), and on malformed input.
Host code only: no device work."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = {"profiles": os.path.join(HERE, "golden", "obs_records_profiles.bin"),
       "surface": os.path.join(HERE, "golden", "obs_records_surface.bin")}
VARS = {"profiles": [1, 2, 3], "surface": [0]}


def restate(rec, var_index, noise):
    """read_obs_mod.F90:400-442 over the records, one row per (record, var)."""
    f32 = np.float32
    sites, rows = [], []
    for i in range(rec.n_records):
        if not rec.multi_level or rec.lev[i] == f32(1.0):
            sites.append((float(rec.lat[i]), float(rec.lon[i]), len(rows)))
        for n in range(rec.nvar):
            v, oe = f32(rec.val[i, 2 * n]), f32(rec.val[i, 2 * n + 1])
            z = f32(0.0) if noise is None else f32(noise[i, n])
            rows.append((var_index[n], float(rec.val[i, 2 * rec.nvar]), float(f32(v + z * oe)),
                         float(f32(f32(1.0) / (oe * oe)))))
    return sites, rows


@pytest.mark.parametrize("name", sorted(FIX))
@pytest.mark.parametrize("with_noise", [False, True])
def test_records_match_restatement(name, with_noise):
    from dida_letkf import obs_records
    rec = obs_records.read_records(FIX[name])
    assert rec.multi_level == (name == "profiles")
    noise = None
    if with_noise:
        noise = np.random.default_rng(5).standard_normal((rec.n_records, rec.nvar)).astype(np.float32)
    out = obs_records.to_rows(rec, VARS[name], noise)
    sites, rows = restate(rec, VARS[name], noise)
    assert out["site_lat"].size == len(sites)
    assert np.array_equal(out["site_lat"], [s[0] for s in sites])
    assert np.array_equal(out["site_lon"], [s[1] for s in sites])
    assert np.array_equal(out["site_row_ptr"], [s[2] for s in sites] + [len(rows)])
    r = np.array(rows)
    assert np.array_equal(out["row_var"], r[:, 0].astype(np.int32))
    assert np.array_equal(out["row_ph"], r[:, 1])
    assert np.array_equal(out["yo"], r[:, 2])
    assert np.array_equal(out["rinv"], r[:, 3])


def test_records_closed_forms():
    from dida_letkf import obs_records
    rec = obs_records.read_records(FIX["profiles"])
    out = obs_records.to_rows(rec, VARS["profiles"])
    # 12 stations, one lacking level 3, one with level 1 only: 12 sites, 55 records
    assert out["site_lat"].size == 12 and rec.n_records == 10 * 5 + 4 + 1
    per_site = np.diff(out["site_row_ptr"]) // 3
    assert sorted(per_site.tolist()) == sorted([5] * 10 + [4, 1])
    assert out["yo"].size == 3 * rec.n_records
    oe = rec.val[:, 1:6:2].reshape(-1).astype(np.float64)
    assert np.abs(out["rinv"] * oe ** 2 - 1.0).max() <= 2.0 ** -22
    assert np.array_equal(out["yo"], rec.val[:, 0:6:2].reshape(-1).astype(np.float64))
    # levels of a profile keep their records' order, pressure falling with level
    for s in range(12):
        ph = out["row_ph"][out["site_row_ptr"][s]:out["site_row_ptr"][s + 1]:3]
        assert np.all(np.diff(ph) < 0)


def test_records_file_round_trip(tmp_path):
    from dida_letkf import obs_records
    rec = obs_records.read_records(FIX["surface"])
    p = tmp_path / "again.bin"
    obs_records.write_records(str(p), rec)
    back = obs_records.read_records(str(p))
    for a in ("lon", "lat", "lev", "val"):
        assert np.array_equal(getattr(back, a), getattr(rec, a))
    bad = tmp_path / "short.bin"
    bad.write_bytes(p.read_bytes()[:-4])
    with pytest.raises(ValueError):
        obs_records.read_records(str(bad))


def test_records_malformed_profiles():
    from dida_letkf import obs_records
    from dida_letkf.engine import DidaError
    rec = obs_records.read_records(FIX["profiles"])
    # a profile record before any lev == 1 (read_obs would store it at site 0)
    cut = obs_records.ObsRecords(rec.nvar, True, rec.lon[1:], rec.lat[1:], rec.lev[1:],
                                 rec.val[1:])
    with pytest.raises(DidaError, match="never started"):
        obs_records.to_rows(cut, VARS["profiles"])
    for badlev in (0.0, 2.5):
        lev = rec.lev.copy()
        lev[3] = badlev
        r2 = obs_records.ObsRecords(rec.nvar, True, rec.lon, rec.lat, lev, rec.val)
        with pytest.raises(DidaError, match="level"):
            obs_records.to_rows(r2, VARS["profiles"])
    empty = obs_records.ObsRecords(3, True, *(np.zeros(0, np.float32),) * 3,
                                   np.zeros((0, 7), np.float32))
    out = obs_records.to_rows(empty, VARS["profiles"])
    assert out["site_lat"].size == 0 and out["yo"].size == 0
